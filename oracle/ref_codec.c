/*
 * Golden-vector harness for the reference's halo message formats (test
 * infrastructure only).
 *
 * Linked with the unmodified reference objects comm.o and util.o (built
 * from /root/reference/ref by oracle/Makefile), it defines the reference's
 * globals the way ref/main.c does (MA_MAIN before block.h, comm.h,
 * timer.h), builds two blocks with the reference's ragged cell arrays
 * (array[var][i][j][k], ref/main.c:429-450) and calls the reference's own
 *
 *   pack_face(buf, block, face_case, dir, start, num_comm)    ref/comm.c:245
 *   unpack_face(buf, block, face_case, dir, start, num_comm)  ref/comm.c:993
 *
 * for every direction and face case (0-9 and 10-19) of one --code.
 *
 * Usage: ref_codec.x nx ny nz code out.bin
 *
 * Cells and the buffers handed to unpack_face come from a 64-bit LCG
 * (seed 12345; value = (x >> 11) * 2^-53), drawn in this order: block 0's
 * cells, block 1's cells (both var, i, j, k order, ghosts included), then
 * per (dir, case) one unpack buffer of MAXBUF doubles.
 *
 * Output (little-endian): int32 header {magic 'AMC1' 0x414d4331, nx, ny,
 * nz, code, num_vars, start, num_comm}, then per dir 0..2, per case in
 * 0..9, 10..19:
 *   int32 packed length L (doubles pack_face wrote, found from a sentinel
 *   fill), then L doubles;
 *   the unpacked block's ghost plane of that direction on the case's side
 *   for variables start..start+num_comm-1: num_comm x P doubles, P the
 *   plane's (n+2)^2 cells in (slower, faster) axis order;
 *   uint64 FNV-1a hash of the unpacked block's whole array (all
 *   variables, ghosts included, var-i-j-k order), which pins every cell
 *   unpack_face did not write as unchanged.
 * The unpacked block is restored to its initial cells before each case.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>

#define MA_MAIN
#include "block.h"
#include "comm.h"
#include "timer.h"
#include "proto.h"

#define NUM_VARS 4
#define START 1
#define NUM_COMM 2
#define MAXBUF 4096

static uint64_t lcg;

static double draw(void)
{
   lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
   return (double) (lcg >> 11) * (1.0 / 9007199254740992.0);
}

static double ****alloc_array(int gx, int gy, int gz)
{
   double ****a = malloc(sizeof(double ***) * NUM_VARS);
   for (int v = 0; v < NUM_VARS; v++) {
      a[v] = malloc(sizeof(double **) * gx);
      for (int i = 0; i < gx; i++) {
         a[v][i] = malloc(sizeof(double *) * gy);
         for (int j = 0; j < gy; j++)
            a[v][i][j] = malloc(sizeof(double) * gz);
      }
   }
   return a;
}

int main(int argc, char **argv)
{
   if (argc != 6) {
      fprintf(stderr, "usage: %s nx ny nz code out.bin\n", argv[0]);
      return 1;
   }
   x_block_size = atoi(argv[1]);
   y_block_size = atoi(argv[2]);
   z_block_size = atoi(argv[3]);
   code = atoi(argv[4]);
   x_block_half = x_block_size / 2;
   y_block_half = y_block_size / 2;
   z_block_half = z_block_size / 2;
   num_vars = NUM_VARS;
   const int gx = x_block_size + 2, gy = y_block_size + 2, gz = z_block_size + 2;
   const int n[3] = { x_block_size, y_block_size, z_block_size };

   blocks = calloc(2, sizeof(block));
   lcg = 12345;
   double ****init[2];
   for (int b = 0; b < 2; b++) {
      blocks[b].array = alloc_array(gx, gy, gz);
      init[b] = alloc_array(gx, gy, gz);
      for (int v = 0; v < NUM_VARS; v++)
         for (int i = 0; i < gx; i++)
            for (int j = 0; j < gy; j++)
               for (int k = 0; k < gz; k++)
                  blocks[b].array[v][i][j][k] = init[b][v][i][j][k] = draw();
   }

   FILE *f = fopen(argv[5], "wb");
   if (!f) { perror(argv[5]); return 1; }
   int32_t hdr[8] = { 0x414d4331, x_block_size, y_block_size, z_block_size, code, NUM_VARS,
                      START, NUM_COMM };
   fwrite(hdr, sizeof(hdr), 1, f);
   static double buf[MAXBUF];
   for (int dir = 0; dir < 3; dir++)
      for (int c = 0; c < 20; c++) {
         /* pack from block 0 */
         for (int t = 0; t < MAXBUF; t++)
            buf[t] = -1.0e300;
         pack_face(buf, 0, c, dir, START, NUM_COMM);
         int32_t len = 0;
         for (int t = 0; t < MAXBUF; t++)
            if (buf[t] != -1.0e300)
               len = t + 1;
         fwrite(&len, sizeof(len), 1, f);
         fwrite(buf, sizeof(double), len, f);

         /* unpack into block 1 */
         for (int v = 0; v < NUM_VARS; v++)
            for (int i = 0; i < gx; i++)
               for (int j = 0; j < gy; j++)
                  memcpy(blocks[1].array[v][i][j], init[1][v][i][j], sizeof(double) * gz);
         for (int t = 0; t < MAXBUF; t++)
            buf[t] = draw();
         unpack_face(buf, 1, c, dir, START, NUM_COMM);
         const int q = c >= 10 ? n[dir] + 1 : 0;
         for (int v = START; v < START + NUM_COMM; v++)
            for (int a = 0; a < (dir == 0 ? gy : gx); a++)
               for (int b = 0; b < (dir == 2 ? gy : gz); b++) {
                  double x = dir == 0 ? blocks[1].array[v][q][a][b]
                           : dir == 1 ? blocks[1].array[v][a][q][b]
                                      : blocks[1].array[v][a][b][q];
                  fwrite(&x, sizeof(x), 1, f);
               }
         uint64_t h = 1469598103934665603ULL;
         for (int v = 0; v < NUM_VARS; v++)
            for (int i = 0; i < gx; i++)
               for (int j = 0; j < gy; j++)
                  for (int k = 0; k < gz; k++) {
                     const unsigned char *p = (const unsigned char *) &blocks[1].array[v][i][j][k];
                     for (int t = 0; t < 8; t++)
                        h = (h ^ p[t]) * 1099511628211ULL;
                  }
         fwrite(&h, sizeof(h), 1, f);
      }
   fclose(f);
   return 0;
}
