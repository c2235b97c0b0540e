/*
 * Golden-vector harness for the reference build (test infrastructure only).
 *
 * Linked into the unmodified reference objects (ref/*.c) by oracle/Makefile.
 * It interposes MPI_Allreduce and forwards every call to PMPI_Allreduce, so
 * the reference computes exactly what it always does; on the side it records
 *
 *   AMR_DUMP_SUMS=<file>   every 1-double MPI_SUM reduction, i.e. each
 *                          check_sum() result (ref/check_sum.c:57), printed
 *                          at %.17g, one per line, in call order;
 *   AMR_DUMP_CELLS=<file>  at the first 142-double reduction, which is the
 *                          first call inside calculate_results
 *                          (ref/profile.c:1409) and runs before deallocate
 *                          (ref/main.c:353), a binary dump of every active
 *                          block in sorted-list order (ghost cells included).
 *
 *   AMR_DUMP_BSUMS=<file>  at the same point, per active block and variable
 *                          the sequential (i,j,k) sum of the interior cells
 *                          and of the ghost cells (meshes too big to dump).
 *
 *   AMR_DUMP_SAMPLE=<file> at the same point, 16384 cells drawn from a
 *                          splitmix64 sequence (seed 0x5eed5eed5eed): per
 *                          cell the sorted-list index, variable, i, j, k
 *                          (ghosts included) and value.  The product's
 *                          driver writes the same draws (host/driver.c).
 *
 * Cell dump format (little-endian):
 *   int32 magic 0x414d5231 ('AMR1'), nx, ny, nz, num_vars, n_active,
 *         num_refine, pad
 *   per block: int64 number, int32 level, int32 slot,
 *              double cells[num_vars][nx+2][ny+2][nz+2]
 * Block-sum format: same header with magic 0x414d5332 ('AMS2'), then per
 *   block: int64 number, int32 level, int32 slot,
 *          double sums[num_vars][2]   (interior, ghost)
 * Sample format: same header with magic 0x414d5333 ('AMS3') and the cell
 *   count in the last word, then per cell: int32 in, var, i, j, k, pad;
 *   double value.
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <mpi.h>

#include "block.h"

static void dump_cells(const char *path)
{
   FILE *f = fopen(path, "wb");
   if (!f) { perror(path); return; }
   int32_t hdr[8];
   int n_active = sorted_index[num_refine + 1];
   hdr[0] = 0x414d5231; hdr[1] = x_block_size; hdr[2] = y_block_size;
   hdr[3] = z_block_size; hdr[4] = num_vars; hdr[5] = n_active;
   hdr[6] = num_refine; hdr[7] = 0;
   fwrite(hdr, sizeof(hdr), 1, f);
   for (int in = 0; in < n_active; in++) {
      int slot = sorted_list[in].n;
      block *bp = &blocks[slot];
      int64_t number = (int64_t) bp->number;
      int32_t ls[2] = { bp->level, slot };
      fwrite(&number, sizeof(number), 1, f);
      fwrite(ls, sizeof(ls), 1, f);
      for (int v = 0; v < num_vars; v++)
         for (int i = 0; i < x_block_size + 2; i++)
            for (int j = 0; j < y_block_size + 2; j++)
               fwrite(bp->array[v][i][j], sizeof(double), z_block_size + 2, f);
   }
   fclose(f);
}

static void dump_block_sums(const char *path)
{
   FILE *f = fopen(path, "wb");
   if (!f) { perror(path); return; }
   int32_t hdr[8];
   int n_active = sorted_index[num_refine + 1];
   hdr[0] = 0x414d5332; hdr[1] = x_block_size; hdr[2] = y_block_size;
   hdr[3] = z_block_size; hdr[4] = num_vars; hdr[5] = n_active;
   hdr[6] = num_refine; hdr[7] = 0;
   fwrite(hdr, sizeof(hdr), 1, f);
   for (int in = 0; in < n_active; in++) {
      int slot = sorted_list[in].n;
      block *bp = &blocks[slot];
      int64_t number = (int64_t) bp->number;
      int32_t ls[2] = { bp->level, slot };
      fwrite(&number, sizeof(number), 1, f);
      fwrite(ls, sizeof(ls), 1, f);
      for (int v = 0; v < num_vars; v++) {
         double s[2] = { 0.0, 0.0 };
         for (int i = 0; i < x_block_size + 2; i++)
            for (int j = 0; j < y_block_size + 2; j++)
               for (int k = 0; k < z_block_size + 2; k++) {
                  int inner = i >= 1 && i <= x_block_size &&
                              j >= 1 && j <= y_block_size &&
                              k >= 1 && k <= z_block_size;
                  s[inner ? 0 : 1] += bp->array[v][i][j][k];
               }
         fwrite(s, sizeof(s), 1, f);
      }
   }
   fclose(f);
}

#define SAMPLE_CELLS 16384

static uint64_t splitmix64(uint64_t *x)
{
   uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

static void dump_sample(const char *path)
{
   FILE *f = fopen(path, "wb");
   if (!f) { perror(path); return; }
   int32_t hdr[8];
   int n_active = sorted_index[num_refine + 1];
   hdr[0] = 0x414d5333; hdr[1] = x_block_size; hdr[2] = y_block_size;
   hdr[3] = z_block_size; hdr[4] = num_vars; hdr[5] = n_active;
   hdr[6] = num_refine; hdr[7] = n_active > 0 ? SAMPLE_CELLS : 0;
   fwrite(hdr, sizeof(hdr), 1, f);
   uint64_t x = 0x5eed5eed5eedULL;
   for (int s = 0; s < hdr[7]; s++) {
      int32_t r[6];
      r[0] = (int32_t) (splitmix64(&x) % (uint64_t) n_active);
      r[1] = (int32_t) (splitmix64(&x) % (uint64_t) num_vars);
      r[2] = (int32_t) (splitmix64(&x) % (uint64_t) (x_block_size + 2));
      r[3] = (int32_t) (splitmix64(&x) % (uint64_t) (y_block_size + 2));
      r[4] = (int32_t) (splitmix64(&x) % (uint64_t) (z_block_size + 2));
      r[5] = 0;
      double v = blocks[sorted_list[r[0]].n].array[r[1]][r[2]][r[3]][r[4]];
      fwrite(r, sizeof(r), 1, f);
      fwrite(&v, sizeof(v), 1, f);
   }
   fclose(f);
}

int MPI_Allreduce(const void *sendbuf, void *recvbuf, int count,
                  MPI_Datatype datatype, MPI_Op op, MPI_Comm comm)
{
   static FILE *sums = NULL;
   static int sums_opened = 0, cells_done = 0;
   int rc = PMPI_Allreduce(sendbuf, recvbuf, count, datatype, op, comm);

   if (count == 1 && datatype == MPI_DOUBLE && op == MPI_SUM) {
      if (!sums_opened) {
         const char *p = getenv("AMR_DUMP_SUMS");
         sums_opened = 1;
         if (p) sums = fopen(p, "w");
      }
      if (sums) {
         fprintf(sums, "%.17g\n", *(const double *) recvbuf);
         fflush(sums);
      }
   } else if (count == 142 && datatype == MPI_DOUBLE && !cells_done) {
      const char *p = getenv("AMR_DUMP_CELLS");
      const char *q = getenv("AMR_DUMP_BSUMS");
      const char *r = getenv("AMR_DUMP_SAMPLE");
      cells_done = 1;
      if (p) dump_cells(p);
      if (q) dump_block_sums(q);
      if (r) dump_sample(r);
   }
   return rc;
}
