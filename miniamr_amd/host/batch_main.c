/*
 * batch_main.c -- miniAMR_batch.x: many independent runs on one GPU.
 *
 *   miniAMR_batch.x [--no_autotune] [--shape vpw,sw,ew,wc,xv,xt] [--quiet] <runs file>
 *
 * --shape fixes the launch shape (amr_batch_shape: variables per stencil
 * item, stencil and exchange grid sizes, workgroup-mode classes, variables
 * and threads per exchange item) instead of autotuning it.
 *
 * Each non-empty line of the runs file not starting with '#' holds one
 * run's command line in the reference's syntax (the same flags as
 * miniAMR.x, plus --seed and --precision).  Every run prints its own
 * Summary line; a last line gives the batch's throughput.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amr.h"
#include "amr_batch.h"
#include "amr_run.h"

#define MAX_TOK 4096

int main(int argc, char **argv)
{
   int autotune = 1, quiet = 0, shape_set = 0;
   amr_batch_shape shp = {1, 0, 0, 0, 8, 128};
   const char *path = NULL;
   for (int i = 1; i < argc; i++) {
      if (!strcmp(argv[i], "--no_autotune"))
         autotune = 0;
      else if (!strcmp(argv[i], "--shape") && i + 1 < argc) {
         if (sscanf(argv[++i], "%d,%d,%d,%d,%d,%d", &shp.vpw, &shp.stencil_wgs,
                    &shp.exchange_wgs, &shp.wg_classes, &shp.exchange_vars,
                    &shp.exchange_threads) < 4) {
            printf("--shape takes vpw,stencil_wgs,exchange_wgs,wg_classes"
                   "[,exchange_vars[,exchange_threads]]\n");
            return -1;
         }
         shape_set = 1;
         autotune = 0;
      } else if (!strcmp(argv[i], "--quiet"))
         quiet = 1;
      else
         path = argv[i];
   }
   if (!path) {
      printf("usage: %s [--no_autotune] [--shape vpw,sw,ew,wc,xv,xt] [--quiet] <runs file>\n",
             argv[0]);
      return -1;
   }
   FILE *f = fopen(path, "r");
   if (!f) {
      printf("cannot open %s\n", path);
      return -1;
   }
   int n = 0, cap = 0;
   int *argcs = NULL;
   char ***argvs = NULL;
   char line[65536];
   while (fgets(line, sizeof(line), f)) {
      char *s = line;
      while (*s == ' ' || *s == '\t')
         s++;
      if (*s == '#' || *s == '\n' || !*s)
         continue;
      if (n == cap) {
         cap = cap ? 2 * cap : 64;
         argcs = (int *) realloc(argcs, sizeof(int) * cap);
         argvs = (char ***) realloc(argvs, sizeof(char **) * cap);
      }
      char **av = (char **) amr_xmalloc(sizeof(char *) * MAX_TOK);
      int ac = 0;
      av[ac++] = argv[0];
      for (char *t = strtok(s, " \t\r\n"); t && ac < MAX_TOK - 1; t = strtok(NULL, " \t\r\n"))
         av[ac++] = strdup(t);
      av[ac] = NULL;
      argcs[n] = ac;
      argvs[n] = av;
      n++;
   }
   fclose(f);
   if (!n) {
      printf("no runs in %s\n", path);
      return -1;
   }
   if (quiet)
      for (int i = 0; i < n; i++) {
         /* drop the per-level block counts and timestep lines */
         char **av = argvs[i];
         av[argcs[i]++] = strdup("--report_perf");
         av[argcs[i]++] = strdup("0");
         av[argcs[i]] = NULL;
      }

   double t0 = amr_now();
   amr_batch_run *B = amr_batch_run_create(n, argcs, argvs, autotune);
   if (shape_set && amr_batch_set_shape(amr_batch_run_batch(B), &shp)) {
      printf("%s\n", amr_batch_last_error(amr_batch_run_batch(B)));
      return -1;
   }
   double t1 = amr_now();
   int rc;
   while (!(rc = amr_batch_run_step(B)))
      ;
   double t2 = amr_now();
   amr_batch_run_finish(B);
   double updates = 0.0;
   for (int i = 0; i < n; i++) {
      amr_run_stats st;
      amr_batch_run_stats_get(B, i, &st);
      updates += st.cell_var_updates;
      printf("run %d: ", i);
      amr_run_report(amr_batch_run_get(B, i));
   }
   amr_batch_shape sh;
   amr_batch_get_shape(amr_batch_run_batch(B), &sh);
   printf("Batch: runs %d setup %lf time %lf device %lf updates %.6e GCell-var/s %.4lf "
          "shape vpw %d stencil_wgs %d exchange_wgs %d wg_classes %d exchange_vars %d "
          "exchange_threads %d tune_ms %.4lf\n",
          n, t1 - t0, t2 - t1, amr_batch_run_device_ms(B) * 1e-3, updates,
          t2 > t1 ? updates / (t2 - t1) * 1e-9 : 0.0, sh.vpw, sh.stencil_wgs, sh.exchange_wgs,
          sh.wg_classes, sh.exchange_vars, sh.exchange_threads, amr_batch_run_tune_ms(B));
   amr_batch_run_destroy(B);
   return rc == 2;
}
