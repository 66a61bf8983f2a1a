/*
 * tester_main.c -- the `tester` battery (src/4_driver/tester.c:28-599).
 *
 * Each test runs two configurations through Runner_run_case and passes iff
 * both give normsqdiff == 0 and equal normsq, compared exactly
 * (compare_runs, src/4_driver/runner.c:187-241).  The reference compiles a
 * different battery per build (serial, OpenMP threads, OpenMP tasks, CUDA,
 * MPI, variants); this GPU build accepts every flag those builds accept, so
 * all batteries run here.  The MPI battery's process-grid pairs
 * (--nproc_x/y > 1) cannot run on one process; its z-blocking and face-comm
 * pairs run with one process.
 *
 * usage: tester [--nm N] [--battery serial|openmp|tasks|cuda|mpi|variants|all]
 *   --nm defaults to 4, the moment count the reference's test scripts build
 *   with (scripts/cmake_serial.sh:21-23); variants use 16 as
 *   scripts/test.bash:321-337 does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "runner.h"

enum { MAX_LINE = 1024 };

static int g_ntest = 0, g_npass = 0, g_nm = 4;

static void pair(Env* env, const char* common, const char* s1, const char* s2) {
  char a[MAX_LINE], b[MAX_LINE];
  snprintf(a, sizeof(a), "--nm %d %s %s", g_nm, common, s1);
  snprintf(b, sizeof(b), "--nm %d %s %s", g_nm, common, s2);
  g_npass += compare_runs(a, b, env) ? 1 : 0;
  g_ntest++;
}

/* tester.c:46-98 */
static void test_serial(Env* env) {
  int key, sx, sy, sz;
  for (key = 0; key < 2; ++key) {
    char common[MAX_LINE], s2[MAX_LINE];
    snprintf(common, sizeof(common),
             "--ncell_x %i --ncell_y %i --ncell_z %i --ne 3 --na 7", 2 + 3 * key,
             1 + 3 * key, 2 + 3 * key);
    for (sx = 1; sx <= 3; ++sx)
      for (sy = 1; sy <= 3; ++sy)
        for (sz = 1; sz <= 3; ++sz) {
          snprintf(s2, sizeof(s2),
                   "--ncell_x_per_subblock %i --ncell_y_per_subblock %i "
                   "--ncell_z_per_subblock %i", sx, sy, sz);
          pair(env, common, "", s2);
        }
  }
}

static void subblock_thread_grid(Env* env, const char* common, const char* extra) {
  int sx, sy, sz, ny, nz, ne, no;
  char s2[MAX_LINE];
  for (sx = 1; sx <= 4; ++sx)
    for (sy = 1; sy <= 5; ++sy)
      for (sz = 1; sz <= 3; ++sz)
        for (ny = 1; ny <= 2; ++ny)
          for (nz = 1; nz <= 2; ++nz)
            for (ne = 1; ne <= 2; ++ne)
              for (no = 1; no <= 2; ++no) {
                snprintf(s2, sizeof(s2),
                         "--ncell_x_per_subblock %i --ncell_y_per_subblock %i "
                         "--ncell_z_per_subblock %i --nthread_y %i --nthread_z %i "
                         "--nthread_e %i --nthread_octant %i %s",
                         sx, sy, sz, ny, nz, ne, no, extra);
                pair(env, common, "", s2);
              }
}

/* tester.c:103-211 */
static void test_openmp(Env* env) {
  const char* pat1 = "--ncell_x 1 --ncell_y 2 --ncell_z 1 --ne 1 --na 1 "
                     "--ncell_x_per_subblock 1 --ncell_y_per_subblock 1 "
                     "--ncell_z_per_subblock 1 --nthread_y %i";
  char s1[MAX_LINE], s2[MAX_LINE];
  int e, k;
  snprintf(s1, sizeof(s1), pat1, 1);
  snprintf(s2, sizeof(s2), pat1, 2);
  pair(env, "", s1, s2);
  for (e = 1; e <= 5; ++e)
    for (k = 0; k <= 3; ++k) {
      const char* pat2 = "--ncell_x 5 --ncell_y 4 --ncell_z 5 --ne 17 --na 10 "
                         "--nthread_e %i --nthread_octant %i";
      snprintf(s1, sizeof(s1), pat2, 1, 1);
      snprintf(s2, sizeof(s2), pat2, e, 1 << k);
      pair(env, "", s1, s2);
    }
  subblock_thread_grid(env, "--ncell_x 3 --ncell_y 4 --ncell_z 2 --ne 2 --na 1", "");
}

/* tester.c:216-285 */
static void test_openmp_tasks(Env* env) {
  const char* s1 = "--ncell_x 9 --ncell_y 13 --ncell_z 3 --ne 5 --na 4 --nblock_z 1";
  int e, k, kx, ky, kz;
  char s2[MAX_LINE];
  for (e = 1; e <= 2; ++e)
    for (k = 0; k <= 3; ++k)
      for (kx = 0; kx < 2; ++kx)
        for (ky = 0; ky < 2; ++ky)
          for (kz = 0; kz < 2; ++kz) {
            snprintf(s2, sizeof(s2),
                     "%s --nthread_e %i --nthread_octant %i --ncell_x_per_subblock %i "
                     "--ncell_y_per_subblock %i --ncell_z_per_subblock %i",
                     s1, e, 1 << k, kx ? 3 : 1, ky ? 5 : 1, kz ? 2 : 1);
            pair(env, "", s1, s2);
          }
}

/* tester.c:290-390 */
static void test_cuda(Env* env) {
  const char* c1 = "--ncell_x 2 --ncell_y 3 --ncell_z 4 --ne 20 --na 5 --nblock_z 2";
  char s2[MAX_LINE];
  int k, e;
  for (k = 0; k <= 3; ++k) {
    snprintf(s2, sizeof(s2), "--is_using_device 1 --nthread_e %i --nthread_octant %i",
             1, 1 << k);
    pair(env, c1, "", s2);
  }
  for (e = 1; e <= 20; ++e) {
    snprintf(s2, sizeof(s2), "--is_using_device 1 --nthread_e %i --nthread_octant %i",
             e, 8);
    pair(env, c1, "", s2);
  }
  subblock_thread_grid(env, "--ncell_x 3 --ncell_y 4 --ncell_z 2 --ne 2 --na 32",
                       "--is_using_device 1");
}

/* tester.c:395-453, one-process pairs */
static void test_mpi(Env* env) {
  pair(env, "--ncell_x 5 --ncell_y 4 --ncell_z 6 --ne 7 --na 10",
       "--nproc_x 1 --nproc_y 1 --nblock_z 1", "--nproc_x 1 --nproc_y 1 --nblock_z 2");
  pair(env, "--ncell_x 5 --ncell_y 4 --ncell_z 6 --ne 7 --na 10 --is_face_comm_async 0",
       "--nblock_z 1", "--nblock_z 2");
  pair(env, "--ncell_x 5 --ncell_y 8 --ncell_z 16 --ne 9 --na 12", "--nblock_z 1",
       "--nblock_z 2");
  pair(env, "--ncell_x 5 --ncell_y 8 --ncell_z 16 --ne 9 --na 12", "--nblock_z 2",
       "--nblock_z 4");
}

/* tester.c:510-546 */
static void test_variants(Env* env) {
  const int nm_save = g_nm;
  int nit;
  char s2[MAX_LINE];
  g_nm = 16;
  for (nit = 1; nit <= 2; ++nit) {
    snprintf(s2, sizeof(s2), "--niterations %i", nit);
    pair(env, "--ncell_x 4 --ncell_y 3 --ncell_z 5 --ne 11 --na 7", "", s2);
  }
  g_nm = nm_save;
}

int main(int argc, char** argv) {
  Env env = Env_null();
  const char* battery = "all";
  int i;
  for (i = 1; i + 1 < argc; i += 2) {
    if (strcmp(argv[i], "--nm") == 0) g_nm = atoi(argv[i + 1]);
    else if (strcmp(argv[i], "--battery") == 0) battery = argv[i + 1];
    else {
      fprintf(stderr, "usage: %s [--nm N] [--battery NAME]\n", argv[0]);
      return 2;
    }
  }
  Env_initialize(&env, argc, argv);
#define RUN(name, fn) \
  if (strcmp(battery, "all") == 0 || strcmp(battery, name) == 0) fn(&env)
  RUN("serial", test_serial);
  RUN("openmp", test_openmp);
  RUN("tasks", test_openmp_tasks);
  RUN("cuda", test_cuda);
  RUN("mpi", test_mpi);
  RUN("variants", test_variants);
#undef RUN
  printf("TESTS %i    PASSED %i    FAILED %i\n", g_ntest, g_npass, g_ntest - g_npass);
  Env_finalize(&env);
  return g_ntest == g_npass ? 0 : 1;
}
