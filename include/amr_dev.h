/*
 * amr_dev.h -- C ABI of the MI355X (gfx950) miniAMR hot path.
 *
 * The reference program has no plugin or FFI layer: its per-timestep data
 * path is five void functions over global state, declared in ref/proto.h.
 * This header is the drop-in seam that replaces them.  The caller (the C
 * driver in miniamr_amd/host, or ctypes from Python) owns the block/object
 * model, the refinement decisions and the slot allocator; the library owns
 * the cell values, which stay resident in device memory for the lifetime of
 * the context.
 *
 *   reference call site                         replaced by
 *   -----------------------------------------   --------------------------
 *   allocate()          ref/main.c:418-605      amr_dev_create
 *   deallocate()        ref/main.c:609-624      amr_dev_destroy
 *   value fill in init  ref/init.c:484-495      amr_dev_upload_block
 *   comm(start,n,stage) ref/comm.c:42-242       amr_dev_set_mesh + amr_dev_comm
 *     on_proc_comm      ref/comm.c:1465-1580      (face jobs, kind SAME)
 *     on_proc_comm_diff ref/comm.c:1585-1908      (kinds PROLONG, RESTRICT)
 *     apply_bc          ref/comm.c:1911-1965      (kind BC)
 *   stencil_driver(var,stage) ref/stencil.c:43-981  amr_dev_stencil_stage
 *                                                 (amr_dev_stencil: stage 0;
 *                                                 --stencil 0 coefficients of
 *                                                 ref/init.c:418-423:
 *                                                 amr_dev_set_materials)
 *   check_sum(var)      ref/check_sum.c:36-65    amr_dev_check_sum
 *   split_blocks() data loop ref/block.c:161-173   amr_dev_split
 *   consolidate_blocks() data loop ref/block.c:411-431  amr_dev_consolidate
 *   reset_all + check_objects + refine_level     amr_dev_mark_refine
 *                       ref/refine.c:66-82, 325-527, ref/move.c:65-623
 *   pack_face / unpack_face (off-rank faces)   amr_dev_pack_faces /
 *                       ref/comm.c:245-1461          amr_dev_unpack_faces
 *
 * Conventions
 *   - Cell arrays use the reference's index order array[var][i][j][k]
 *     (ref/block.h:52): i = x, j = y, k = z (fastest), one ghost layer,
 *     interior indices 1..n.
 *   - Faces are numbered 0..5 = W, E, S, N, D, U (ref/block.h:49).
 *   - Every function returns AMR_OK (0) or a negative AMR_ERR_* code; the
 *     text of the last error is available from amr_dev_last_error.
 *   - One HIP stream per context; calls on one context must not overlap
 *     (the reference's functions are not reentrant either).  Functions that
 *     return values to the host (check_sum, download) synchronise the
 *     stream; the others only enqueue work.
 */
#ifndef AMR_DEV_H
#define AMR_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct amr_dev amr_dev;

enum {
   AMR_OK = 0,
   AMR_ERR_ARG = -1,     /* bad argument (size, slot, var out of range) */
   AMR_ERR_HIP = -2,     /* HIP runtime error */
   AMR_ERR_NOMEM = -3,   /* device allocation failed */
   AMR_ERR_STATE = -4,   /* call out of order (e.g. comm before set_mesh) */
   AMR_ERR_NODEV = -5    /* no usable gfx950 device */
};

/* Storage precision of the cell values. */
enum {
   AMR_PREC_F64 = 0,     /* reference arithmetic, bit-exact */
   AMR_PREC_F32 = 1,     /* FP32 storage, FP32 sums, FP64 checksums */
   AMR_PREC_BF16 = 2     /* BF16 storage, FP32 sums, FP64 checksums */
};

/* Face-job kinds of one exchange direction.  Kinds 0-3 are the --code 0
 * exchange (ghost cells of the face interior; the 27-point stencil extends
 * same-level y/z faces and boundary faces over the ghost ring).  Kinds 4-7
 * (kind | AMR_FACE_GHOSTS) are the --code 1/2 exchange, which moves whole
 * (n+2) x (n+2) ghost planes, ghost ring included, at any stencil
 * (ref/comm.c:1535-1579, 1689-1907, 1941-1964); on one rank codes 1 and 2
 * differ only in their message formats, so both use these kinds. */
enum {
   AMR_FACE_SAME = 0,     /* dst ghost plane <- src boundary plane
                             (ref/comm.c:1473-1534) */
   AMR_FACE_PROLONG = 1,  /* dst is fine: ghost <- coarse cell / 4
                             (ref/comm.c:1618-1625 and the y/z twins) */
   AMR_FACE_RESTRICT = 2, /* dst is coarse: ghost quarter <- sum of 4 fine
                             cells (ref/comm.c:1626-1630 and twins) */
   AMR_FACE_BC = 3,       /* reflective boundary (ref/comm.c:1911-1965) */
   AMR_FACE_GHOSTS = 4,   /* flag: whole ghost planes (--code 1/2) */
   AMR_FACE_SAME_GHOSTS = 4,     /* (n+2)^2 plane copy (ref/comm.c:1535-1579) */
   AMR_FACE_PROLONG_GHOSTS = 5,  /* fine ghost plane, ring included: fine
                                    (a, b) <- coarse (a0 + (a+1)/2,
                                    b0 + (b+1)/2) / 4 (ref/comm.c:1690-1761,
                                    y and z at :1762-1906) */
   AMR_FACE_RESTRICT_GHOSTS = 6, /* the coarse quarter and its share of the
                                    ring: interior cells <- 4 fine cells,
                                    ring edges <- 2 fine ring cells, ring
                                    corner <- 1 (same lines) */
   AMR_FACE_BC_GHOSTS = 7        /* reflective boundary over the whole plane */
};

typedef struct {
   int nx, ny, nz;     /* interior block size, each even and > 0 */
   int num_vars;       /* > 0 */
   int max_blocks;     /* number of block slots (ref: --max_blocks) */
   int stencil;        /* 7 or 27, or 0: the variable-work stencils */
   int precision;      /* AMR_PREC_* */
   int device;         /* HIP device ordinal */
} amr_dev_config;

/* One ghost-plane fill.  `face` is the face of `dst` whose ghost plane is
 * written.  SAME/PROLONG/RESTRICT read `src`; BC reads `dst` itself.
 * (iq, jq) is the quarter of the coarse block's face that the fine block
 * covers, as in ref/comm.c:1585 (iq selects the faster in-plane axis). */
typedef struct {
   int32_t kind, face, dst, src, iq, jq;
} amr_face_job;

/* child interior <- parent octant / 8  (ref/block.c:161-173).
 * Octant o: x half o%2, y half (o/2)%2, z half o/4. */
typedef struct {
   int32_t parent, child, octant, pad;
} amr_split_job;

/* dst octant <- sum of each 2x2x2 group of child cells
 * (ref/block.c:411-431), same octant numbering. */
typedef struct {
   int32_t dst, child, octant, pad;
} amr_merge_job;

int amr_dev_create(const amr_dev_config *cfg, amr_dev **out);
int amr_dev_destroy(amr_dev *dev);
const char *amr_dev_last_error(const amr_dev *dev);

/* Copy one slot's cells, [num_vars][nx+2][ny+2][nz+2] doubles, to/from the
 * device (converted to/from the storage precision). */
int amr_dev_upload_block(amr_dev *dev, int slot, const double *cells);
int amr_dev_download_block(amr_dev *dev, int slot, double *cells);

/* Install the active-block list and the exchange plan after each change of
 * the mesh.  sorted_slots[n_active] lists the active slots in the
 * reference's sorted-list order (ref/block.h:72-77), which fixes the
 * checksum summation order.  jobs[] holds dir_count[0] x-direction jobs,
 * then dir_count[1] y-direction jobs, then dir_count[2] z-direction jobs. */
int amr_dev_set_mesh(amr_dev *dev, const int32_t *sorted_slots, int n_active,
                     const amr_face_job *jobs, const int32_t dir_count[3]);

/* Ghost exchange for variables [start, start+num_comm): three ordered
 * phases x, y, z, or permutations[stage % 6] when permute != 0
 * (ref/comm.c:45-55). */
int amr_dev_comm(amr_dev *dev, int start, int num_comm, int stage, int permute);

/* 7- or 27-point averaging stencil of one variable over all active blocks.
 * (A --stencil 0 context: stencil_driver(var, 0), see below.) */
int amr_dev_stencil(amr_dev *dev, int var);

/* --stencil 0, the variable-work stencils (ref/stencil.c:43-73, 147-981).
 * amr_dev_set_materials installs the coefficients the reference draws at
 * initialisation (ref/init.c:418-423): a1 and a0[n], n = num_vars / 4
 * (the number of "materials"); a --stencil 0 context needs it before its
 * first stencil.  amr_dev_stencil_stage is stencil_driver(var, calc_stage):
 * variable 0 and variables >= 4*n take the 7-point average, the others one
 * of six updates coupled to other variables, chosen by calc_stage % 6, and
 * then the clamp of stencil_check; at 7 or 27 it equals amr_dev_stencil.
 * amr_dev_stage passes its stage as the calculation stage and updates the
 * variables of its range in order, as the reference's loop does.
 * amr_dev_read_check_counts returns how many clamps divided (counts[0]:
 * one divide and two adds each) and scaled (counts[1]: one multiply and one
 * add) since the context was created -- the data-dependent operation counts
 * of the report (ref/stencil.c:966-981). */
int amr_dev_set_materials(amr_dev *dev, double a1, const double *a0, int n);
int amr_dev_stencil_stage(amr_dev *dev, int var, int calc_stage);
int amr_dev_read_check_counts(amr_dev *dev, int64_t counts[2]);

/* Sum of one variable's interior cells over all active blocks (FP64). */
int amr_dev_check_sum(amr_dev *dev, int var, double *sum);

/* Data half of split_blocks / consolidate_blocks.  Jobs are applied as if
 * one after another in array order (a slot freed by an earlier job may be
 * the target of a later one). */
int amr_dev_split(amr_dev *dev, const amr_split_job *jobs, int n);
int amr_dev_consolidate(amr_dev *dev, const amr_merge_job *jobs, int n);

/* amr_dev_split without device allocation.  No counterpart in the
 * reference's interface: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same jobs, same cells as amr_dev_split; the job list
 * and the parent-octant snapshots are staged in a static 16 MiB device
 * array per GPU instead of buffers the context grows.  Limits: at most
 * 65,536 jobs per call, and one parent octant of all variables
 * (num_vars * nx*ny*nz/8 cells) must fit 15 MiB (AMR_ERR_ARG otherwise).
 * Refused while the tape records (AMR_ERR_STATE).  Returns after the work
 * is done; calls that overlap on one GPU (from several threads, on any of
 * its contexts) share the array and race. */
int amr_dev_split_noalloc(amr_dev *dev, const amr_split_job *jobs, int n);

/* Halo message codec: the one-GPU part of an exchange across GPUs.  The
 * reference packs every off-rank face into an MPI message and unpacks what
 * it receives (pack_face ref/comm.c:245-990, unpack_face :993-1461,
 * called from comm() :87-94, :147-151; message lengths ref/init.c:77-117);
 * across GPUs the packed buffers would travel by RCCL send/recv.  A
 * message is one face of one block for variables [start, start+num_comm),
 * variable after variable, in the reference's formats for --code 0, 1 or 2
 * (code 0: face interiors, restriction as 2x2 sums on the fine sender,
 * prolongation as value / 4 on the coarse sender replicated by the fine
 * receiver; codes 1 and 2 add the ghost ring, code 2 prolongs on the send
 * side).  face_case is the reference's comm_face_case: 0 whole face, 1
 * whole face with the 27-point ring, 2-5 whole face <-> quarter (this block
 * is the fine one; quarter = face_case - 2), 6-9 quarter <-> whole face
 * (this block is the coarse one), +10 for the E/N/U face of the direction.
 * pack reads the block's boundary plane, unpack writes its ghost plane.
 * Values are in the context's storage type (double at AMR_PREC_F64, as the
 * reference's MPI_DOUBLE messages).  buf is device memory holding buf_elems
 * values; message t occupies [offset, offset + num_comm * len) with len =
 * amr_dev_face_msg_len(dev, code, dir, face_case, unpack).  Messages of one
 * pack call may share a block; those of one unpack call must write
 * different ghost cells.  Enqueue only (the message list is copied first). */
typedef struct {
   int32_t slot;       /* block slot */
   int32_t face_case;  /* 0..19, as above */
   int64_t offset;     /* first value of the message in buf */
} amr_face_msg;

/* Values per variable of one message (ref/init.c:77-117): the sender's
 * length (unpack = 0) or the receiver's (unpack = 1) for this face case;
 * negative AMR_ERR_* on bad arguments. */
int amr_dev_face_msg_len(amr_dev *dev, int code, int dir, int face_case, int unpack);
int amr_dev_pack_faces(amr_dev *dev, int code, int dir, int start, int num_comm,
                       const amr_face_msg *msgs, int n, void *buf, size_t buf_elems);
int amr_dev_unpack_faces(amr_dev *dev, int code, int dir, int start, int num_comm,
                         const amr_face_msg *msgs, int n, const void *buf, size_t buf_elems);

/* Largest |a - b| over the interior cells of the active blocks, for each
 * variable of [start, start+num): two contexts on one device holding the
 * same mesh (block size, variables, slots and active list), in any
 * storage precisions -- the cell error of a narrow storage form against an
 * FP64 run after each checksum interval.  No counterpart in the reference
 * (its cells are FP64 only).  Waits for b's enqueued work; returns after
 * the comparison. */
int amr_dev_compare_cells(amr_dev *a, amr_dev *b, int start, int num, double *max_abs);

/* Device-side checksum table: n_rows rows of num_vars doubles that
 * amr_dev_stage fills, so a run can enqueue many stages and read the sums
 * back once.  Rows are indexed by the caller (e.g. checksum event number
 * within a refinement interval). */
int amr_dev_reserve_sums(amr_dev *dev, int n_rows);
int amr_dev_read_sums(amr_dev *dev, int row0, int n_rows, double *out);

/* One stage for variables [start, start+num): the three exchange phases
 * (as amr_dev_comm) then the stencil of each variable; when sum_row >= 0
 * the checksum of each variable is stored in row sum_row of the checksum
 * table (columns start..start+num-1).  Enqueue only. */
int amr_dev_stage(amr_dev *dev, int start, int num, int stage, int permute,
                  int sum_row);

/* amr_dev_stage with same-level faces read in place (no counterpart in the
 * reference's interface; the drop-in driver uses it for every stage but the
 * last one of a timestep that ends in a refinement or ends the run).
 * Cells, checksums and every ghost value a stencil reads are those of
 * amr_dev_stage.  The difference is where they come from: the ghost planes
 * of same-level faces of blocks away from the domain boundary, and their
 * 27-point edge extensions, are not written -- the stencil reads those
 * cells from the neighbours' interiors, which stay unchanged because the
 * stage writes its results to a second copy of every interior (records
 * hold two unless the context was created with MINIAMR_FUSED=0 in the
 * environment, which costs ~70% more pool memory at 16^3).  Those ghost
 * cells are stale until the next amr_dev_stage or amr_dev_comm of the same
 * variables; set_mesh, split, consolidate, stencil, download_block, the
 * halo codec and save_state first bring them up to date (to the values
 * amr_dev_stage would have written, from the interiors the stage read).
 * It is amr_dev_stage without two interiors, with permute != 0, while the
 * tape records, and for stages of fewer than 1e8 cell updates
 * (latency-bound; MINIAMR_FUSED=2 at creation fuses those too). */
int amr_dev_stage_fused(amr_dev *dev, int start, int num, int stage, int permute,
                        int sum_row);

/* Wait for all enqueued work. */
int amr_dev_sync(amr_dev *dev);

/* The bandwidth a plain streaming kernel reaches on the current device, in
 * GB/s: a copy of `bytes` bytes (a multiple of 4096; one 16-byte
 * nontemporal load and store per lane, one workgroup of 256 threads per
 * 4 KiB tile) and a read of the same buffer, best of `reps` launches each.
 * Allocates and frees two buffers of `bytes`.  No counterpart in the
 * reference: it measures the practical ceiling that the stencil's roofline
 * figure is read against (bench.py reports both). */
int amr_dev_copy_bandwidth(size_t bytes, int reps, double *copy_gbs, double *read_gbs);

/* Device time of the kernels enqueued since the last call, in
 * milliseconds, split as [0] exchange, [1] stencil, [2] checksum,
 * [3] split/consolidate.  Requires amr_dev_set_timing(dev, 1); level 2
 * also times each exchange direction (amr_dev_read_exchange_timing), at
 * the cost of two more events per direction and stage; level 3 times the
 * stencil launches only (two events per stage; the other entries read 0). */
int amr_dev_set_timing(amr_dev *dev, int on);
int amr_dev_read_timing(amr_dev *dev, double ms[4]);
/* The exchange time of ms[0] split by direction (x, y, z), accumulated by
 * amr_dev_read_timing since the last call of this one.  For the end-of-run
 * report's per-direction rows (ref/profile.c:1085-1166); no counterpart
 * in the reference's call interface.  Zero unless timing level 2. */
int amr_dev_read_exchange_timing(amr_dev *dev, double ms[3]);

/* Bytes of device (and pinned host) memory the context holds, and in how
 * many buffers: the report's memory lines (ref/profile.c:1042-1053), whose
 * reference counterpart counts the program's own mallocs. */
int amr_dev_memory(const amr_dev *dev, size_t *bytes, int *buffers);

/* Refinement marking on the device: reset_all, check_objects (or the
 * uniform rule) and the 2:1 balance fixpoint refine_level of one
 * refinement step (ref/refine.c:66-82, 325-527; ref/move.c:65-623), in one
 * launch with no host round trip between the fixpoint's passes.  Input:
 * the active blocks in sorted order (level_start[l] = first block of level
 * l, num_refine + 2 entries), the parent records and the objects.  Output:
 * the refine flag (1 refine, 0 stay, -1 may coarsen) of every block and of
 * every parent record with number >= 0 -- the flags the reference's
 * sequential passes leave.  Returns AMR_ERR_ARG when the flags do not fit
 * the kernel's local memory (n_blocks + n_parents > 65536) or an object
 * type is unknown; the caller then marks on the host. */
typedef struct {
   int32_t type, pad;
   double cen[3], size[3];
} amr_mark_object;

typedef struct {
   int32_t level;
   int32_t parent;        /* index into parents[], -1 for an initial block */
   int32_t cen[3];        /* centre in units of the finest half-cell grid */
   int32_t nei_level[6];  /* level of the face neighbour; -2 on the boundary */
   int32_t nei[6];        /* index in blocks[] of a same-level or coarser
                             face neighbour, else -1 */
} amr_mark_block;

typedef struct {
   int32_t number;        /* < 0: unused record */
   int32_t child[8];      /* index in blocks[], -1 when the child is a parent */
} amr_mark_parent;

typedef struct {
   int32_t num_refine, cur_max_level, uniform_refine, refine_ghost;
   int32_t mesh_size[3];  /* finest half-cell units per axis */
   int32_t n_objects;
} amr_mark_params;

int amr_dev_mark_refine(amr_dev *dev, const amr_mark_params *p, const amr_mark_block *blocks,
                        int n_blocks, const int32_t *level_start, const amr_mark_parent *parents,
                        int n_parents, const amr_mark_object *objects, int8_t *block_refine,
                        int8_t *parent_refine);

#ifdef __cplusplus
}
#endif

#endif /* AMR_DEV_H */
