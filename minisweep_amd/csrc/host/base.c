/*
 * base.c -- Arguments, Env, Insist, Dimensions, Pointer for the drop-in API.
 *
 * Behaviour follows the reference's base layer:
 *   Arguments  src/1_base/arguments.c:37-198 (read-once int flags, atoi)
 *   insist_    src/1_base/env_assert.c:27-32
 *   Env        src/1_base/env.c:35-85 (one process: nproc == 1)
 *   Dimensions src/2_sweeper_base/dimensions.c:36-101
 *   Pointer    src/1_base/pointer.c:36-267 (host mirror pinned on request,
 *              device copy in HBM)
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <hip/hip_runtime_api.h>

#include "sweeper.h"

/* ---- Insist ------------------------------------------------------------- */

void insist_(const char* condition_string, const char* file, int line) {
  fprintf(stderr, "Insist error: \"%s\". At file %s, line %i.\n",
          condition_string, file, line);
  exit(EXIT_FAILURE);
}

static void hip_insist(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    insist_(buf, file, line);
  }
}
#define HIP_INSIST(call) hip_insist((call), #call, __FILE__, __LINE__)

/* ---- Arguments ---------------------------------------------------------- */

Arguments Arguments_null(void) {
  Arguments a;
  memset(&a, 0, sizeof(a));
  return a;
}

void Arguments_create(Arguments* args, int argc, char** argv) {
  int i;
  args->argc = argc;
  args->argstring = NULL;
  args->argv_unconsumed = (char**)malloc((argc > 0 ? argc : 1) * sizeof(char*));
  for (i = 0; i < argc; ++i) args->argv_unconsumed[i] = argv[i];
}

/* Whitespace-separated tokens; element 0 is an empty program name. */
void Arguments_create_from_string(Arguments* args, const char* argstring) {
  const size_t len = strlen(argstring);
  size_t i;
  int prev_delim = 1;
  args->argstring = (char*)malloc(len + 1);
  memcpy(args->argstring, argstring, len + 1);
  args->argv_unconsumed = (char**)malloc((len + 2) * sizeof(char*));
  args->argv_unconsumed[0] = &args->argstring[len];
  args->argc = 1;
  for (i = 0; i < len; ++i) {
    if (args->argstring[i] == ' ' || args->argstring[i] == '\t') args->argstring[i] = 0;
    {
      const int is_delim = args->argstring[i] == 0;
      if (prev_delim && !is_delim) args->argv_unconsumed[args->argc++] = &args->argstring[i];
      prev_delim = is_delim;
    }
  }
}

void Arguments_destroy(Arguments* args) {
  free(args->argv_unconsumed);
  free(args->argstring);
  args->argv_unconsumed = NULL;
  args->argstring = NULL;
  args->argc = 0;
}

Bool_t Arguments_exists(const Arguments* args, const char* arg_name) {
  int i;
  for (i = 0; i < args->argc; ++i)
    if (args->argv_unconsumed[i] && strcmp(args->argv_unconsumed[i], arg_name) == 0)
      return Bool_true;
  return Bool_false;
}

int Arguments_consume_int_(Arguments* args, const char* arg_name) {
  int result = 0, found = 0, i;
  for (i = 0; i < args->argc; ++i) {
    if (!args->argv_unconsumed[i]) continue;
    if (strcmp(args->argv_unconsumed[i], arg_name) == 0) {
      found = 1;
      args->argv_unconsumed[i] = NULL;
      ++i;
      Insist(i < args->argc);
      Insist(args->argv_unconsumed[i] != NULL);
      result = atoi(args->argv_unconsumed[i]);
      args->argv_unconsumed[i] = NULL;
    }
  }
  Insist(found ? "Invalid use of argument." : 0);
  return result;
}

int Arguments_consume_int_or_default(Arguments* args, const char* arg_name,
                                     int default_value) {
  return Arguments_exists(args, arg_name) ? Arguments_consume_int_(args, arg_name)
                                          : default_value;
}

Bool_t Arguments_are_all_consumed(const Arguments* args) {
  int i;
  for (i = 1; i < args->argc; ++i)
    if (args->argv_unconsumed[i] != NULL) return Bool_false;
  return Bool_true;
}

/* ---- Env ---------------------------------------------------------------- */

Env Env_null(void) {
  Env e;
  memset(&e, 0, sizeof(e));
  return e;
}

/* The device stream is created on first use (Env_stream), so argument
 * errors are reported before anything touches the GPU. */
void Env_initialize(Env* env, int argc, char** argv) {
  (void)argc;
  (void)argv;
  env->is_using_device_ = Bool_true;
  env->is_state_resident_ = Bool_false;
  env->precision_ = MSW_PREC_FP64;
}

void* Env_stream(Env* env) {
  if (!env->stream_) {
    hipStream_t s;
    HIP_INSIST(hipStreamCreate(&s));
    env->stream_ = (void*)s;
  }
  return env->stream_;
}

/* --is_using_device (src/1_base/env_cuda.c:87-90) and --nproc_x/y
 * (src/1_base/env_mpi.c:114-117) are accepted; one process means
 * nproc_x == nproc_y == 1.  --precision selects the sweep arithmetic. */
void Env_set_values(Env* env, Arguments* args) {
  const int dev = Arguments_consume_int_or_default(args, "--is_using_device", 1);
  const int npx = Arguments_consume_int_or_default(args, "--nproc_x", 1);
  const int npy = Arguments_consume_int_or_default(args, "--nproc_y", 1);
  Insist(dev == 0 || dev == 1 ? "Invalid is_using_device supplied." : 0);
  Insist(npx == 1 && npy == 1
             ? "This build runs one process on one GPU: nproc_x and nproc_y must be 1." : 0);
  env->is_using_device_ = Bool_true;
  env->precision_ = Arguments_consume_int_or_default(args, "--precision", MSW_PREC_FP64);
  Insist(env->precision_ >= MSW_PREC_FP64 && env->precision_ <= MSW_PREC_MIXED_BF16
             ? "Invalid precision supplied." : 0);
}

void Env_finalize(Env* env) {
  if (env->stream_) {
    (void)hipStreamDestroy((hipStream_t)env->stream_);
    env->stream_ = NULL;
  }
}

Bool_t Env_is_proc_master(Env* env) { (void)env; return Bool_true; }
Bool_t Env_is_proc_active(const Env* env) { (void)env; return Bool_true; }

Timer Env_get_time(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (Timer)tv.tv_sec + (Timer)tv.tv_usec * 1e-6;
}

/* src/1_base/env.c:81-85: barrier then time; here the barrier is the
 * device's queue draining. */
Timer Env_get_synced_time(Env* env) {
  if (env && env->stream_) HIP_INSIST(hipStreamSynchronize((hipStream_t)env->stream_));
  return Env_get_time();
}

int Env_nproc(const Env* env) { (void)env; return 1; }
int Env_nproc_x(const Env* env) { (void)env; return 1; }
int Env_nproc_y(const Env* env) { (void)env; return 1; }
int Env_proc_x_this(const Env* env) { (void)env; return 0; }
int Env_proc_y_this(const Env* env) { (void)env; return 0; }
double Env_sum_d(Env* env, double value) { (void)env; return value; }
P Env_sum_P(Env* env, P value) { (void)env; return value; }
Bool_t Env_cuda_is_using_device(const Env* env) { return env->is_using_device_; }
void Env_set_state_resident(Env* env, Bool_t resident) {
  env->is_state_resident_ = resident;
}

/* ---- Dimensions --------------------------------------------------------- */

size_t Dimensions_size_state(const Dimensions d, int nu) {
  return (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.nm * nu;
}
size_t Dimensions_size_state_angles(const Dimensions d, int nu) {
  return (size_t)d.ncell_x * d.ncell_y * d.ncell_z * d.ne * d.na * nu * NOCTANT;
}
size_t Dimensions_size_facexy(const Dimensions d, int nu, int noct) {
  return (size_t)d.ncell_x * d.ncell_y * d.ne * d.na * nu * noct;
}
size_t Dimensions_size_facexz(const Dimensions d, int nu, int noct) {
  return (size_t)d.ncell_x * d.ncell_z * d.ne * d.na * nu * noct;
}
size_t Dimensions_size_faceyz(const Dimensions d, int nu, int noct) {
  return (size_t)d.ncell_y * d.ncell_z * d.ne * d.na * nu * noct;
}

/* ---- Pointer ------------------------------------------------------------ */

Pointer Pointer_null(void) {
  Pointer p;
  memset(&p, 0, sizeof(p));
  return p;
}

void Pointer_create(Pointer* p, size_t n, Bool_t is_using_device) {
  *p = Pointer_null();
  p->n_ = n;
  p->is_using_device_ = is_using_device;
}

void Pointer_create_alias(Pointer* p, Pointer* source, size_t base, size_t n) {
  Insist(base + n <= source->n_);
  *p = Pointer_null();
  p->n_ = n;
  p->is_using_device_ = source->is_using_device_;
  p->is_pinned_ = source->is_pinned_;
  p->is_alias_ = Bool_true;
  if (source->h_) p->h_ = source->h_ + base;
  if (source->d_) p->d_ = source->d_ + base;
}

void Pointer_set_pinned(Pointer* p, Bool_t is_pinned) { p->is_pinned_ = is_pinned; }

void Pointer_allocate(Pointer* p) {
  Insist(!p->is_alias_);
  if (p->is_pinned_) {
    HIP_INSIST(hipHostMalloc((void**)&p->h_, p->n_ * sizeof(P) + 1, hipHostMallocDefault));
  } else {
    p->h_ = (P*)malloc(p->n_ * sizeof(P) + 1);
    Insist(p->h_ != NULL);
  }
  if (p->is_using_device_) HIP_INSIST(hipMalloc((void**)&p->d_, p->n_ * sizeof(P) + 1));
}

void Pointer_deallocate(Pointer* p) {
  if (p->is_alias_) return;
  if (p->h_) {
    if (p->is_pinned_) (void)hipHostFree(p->h_);
    else free(p->h_);
  }
  if (p->d_) (void)hipFree(p->d_);
  p->h_ = NULL;
  p->d_ = NULL;
}

void Pointer_destroy(Pointer* p) {
  Pointer_deallocate(p);
  *p = Pointer_null();
}

void Pointer_update_h(Pointer* p) {
  if (p->is_using_device_)
    HIP_INSIST(hipMemcpy(p->h_, p->d_, p->n_ * sizeof(P), hipMemcpyDeviceToHost));
}

void Pointer_update_d(Pointer* p) {
  if (p->is_using_device_)
    HIP_INSIST(hipMemcpy(p->d_, p->h_, p->n_ * sizeof(P), hipMemcpyHostToDevice));
}

void Pointer_update_h_stream(Pointer* p, void* stream) {
  if (p->is_using_device_)
    HIP_INSIST(hipMemcpyAsync(p->h_, p->d_, p->n_ * sizeof(P), hipMemcpyDeviceToHost,
                              (hipStream_t)stream));
}

void Pointer_update_d_stream(Pointer* p, void* stream) {
  if (p->is_using_device_)
    HIP_INSIST(hipMemcpyAsync(p->d_, p->h_, p->n_ * sizeof(P), hipMemcpyHostToDevice,
                              (hipStream_t)stream));
}
