/*
 * main.c -- entry point of the drop-in miniAMR (ref/main.c:38-362).
 *
 * Parses the reference's command line, validates it, allocates the block
 * model and the device context, runs the timestep loop and prints the
 * report.  One rank, one GPU: there is no MPI, and the parameters the
 * reference broadcasts to other ranks stay local.
 */
#include <stdlib.h>
#include "amr.h"

int main(int argc, char **argv)
{
   static amr_run run;
   double t0 = amr_now();

   amr_parse_args(&run.P, argc, argv);
   if (amr_check_input(&run.P))
      exit(-1);
   if (!run.P.block_change)
      run.P.block_change = run.P.num_refine;

   amr_allocate(&run);
   run.t_main = amr_now() - t0;
   amr_driver(&run);
   amr_report(&run);
   amr_dump_final(&run);
   amr_free(&run);
   fflush(NULL);
   return 0;
}
