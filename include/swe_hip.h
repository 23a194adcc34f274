/*
 * libswe_hip.so -- MI355X (gfx950) HIP implementation of the RBF-FD
 * shallow-water RK4 hot path, behind a plain C ABI.
 *
 * It replaces, as a drop-in, the two compute back-ends the reference driver
 * selects between at swe_code/main/main.c:338-356:
 *   - the CPU/OpenMP/OpenACC stage driver
 *       void RK_substep(fType* H, fType* K, fType* F, fType* D, int substep_id)
 *       (swe_code/main/include/rk4_rbffd_swe.h:8, rk4_rbffd_swe.c:30-72)
 *   - the OpenCL path: platform/context/queue set-up and kernel compilation
 *     (device_setup.c:13-225), load_all_buffers (buffers.c:26), per-attempt
 *     buffer mapping (main.c:304-322), RK_substep_ocl (RK_ocl.c:9) and the
 *     read-back map (main.c:379-380).
 * The static model data are handed over once, in the caller's host layout
 * (the cfdl/sfdl/SIMD_LENGTH index macros of swe_config.h:21-54), and
 * repacked on the device; the state H stays on the device between calls.
 *
 * Conventions: every function returns 0 (SWE_HIP_OK) or a negative
 * SWE_HIP_E* code; swe_hip_strerror() names the code and
 * swe_hip_last_error() gives the detail of the most recent failure on the
 * calling thread.  The reference's OpenCL helpers print and exit(-1) on any
 * error (device_setup.c:346); a host that wants that behaviour prints
 * swe_hip_last_error() and exits.  Host buffers are caller-owned; device
 * buffers are owned by the context.  One host thread per context; work is
 * ordered on the context's stream and the functions that return data to the
 * host synchronise it.  FP64, or the reference's FP32 model type through
 * SWE_HIP_FP32 contexts.
 */
#ifndef SWE_HIP_H
#define SWE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_HIP_OK 0
#define SWE_HIP_EINVAL (-1)   /* bad argument or unsupported layout         */
#define SWE_HIP_ENOMEM (-2)   /* device or host allocation failed           */
#define SWE_HIP_EDEVICE (-3)  /* no HIP device / HIP runtime error          */
#define SWE_HIP_ERANGE (-4)   /* a stencil index outside the patch          */

/* Context flags. */
#define SWE_HIP_EXACT 1u      /* sequential stencil sums, no FMA contraction:
                                 bitwise equal to the reference CPU build
                                 compiled without contraction              */
#define SWE_HIP_FP32 2u       /* the reference's FP32 model type (RUN_32BIT,
                                 fType = float, swe_config.h:75-80): static
                                 data and state held and computed in float,
                                 update_H's dt/6.0 product in double as the
                                 reference's C promotion does; one state
                                 (members == 1); state I/O through the *32
                                 entry points                              */

/* The static model data of the local patch (PSMD_struct,
 * swe_code/main/include/swe_config.h:187-233) as plain pointers, in the
 * host layout described by cfdl/sfdl/simd.  Serial single patch only:
 * compute_pid_s == 0 and patch_size == compute_size (the whole sphere runs
 * on one GPU; init_patches.c serial branch). */
typedef struct swe_hip_static {
    int compute_size;          /* padded node count                        */
    int patch_size;
    int compute_pid_s;
    int Nnbr;                  /* real stencil width (1..64)               */
    int padded_Nnbr;           /* row stride of idx/Dx/Dy/Dz/L             */
    int cfdl, sfdl, simd;      /* USE_CFDL, USE_SFDL, SIMD_LENGTH          */
    double gh0, dt;
    const double *x, *y, *z, *f, *ghm;         /* [compute_size]           */
    const double *p_u, *p_v, *p_w, *gradghm;   /* VECT_LIN_ID layout       */
    const int *idx;                            /* DM_LIN_ID layout, 0-based*/
    const double *Dx, *Dy, *Dz, *L;            /* DM_LIN_ID layout, scaled */
} swe_hip_static;

/* The same static data in the reference's FP32 model type (its PSMD_struct
 * with fType = float), for swe_hip_create32. */
typedef struct swe_hip_static32 {
    int compute_size, patch_size, compute_pid_s, Nnbr, padded_Nnbr;
    int cfdl, sfdl, simd;
    float gh0, dt;
    const float *x, *y, *z, *f, *ghm;
    const float *p_u, *p_v, *p_w, *gradghm;
    const int *idx;
    const float *Dx, *Dy, *Dz, *L;
} swe_hip_static32;

typedef struct swe_hip_ctx swe_hip_ctx;

/* Per-phase device time accumulated by the drop-in path while timing is on,
 * seconds (the t_eval_rhs / t_eval_K / t_update_D / t_update_H fields of
 * timing_struct, swe_config.h:262-279). */
typedef struct swe_hip_timers {
    double t_eval_rhs, t_eval_K, t_update_D, t_update_H;
} swe_hip_timers;

/* Replaces createPlatform/createContext/createCommandQueue/compileKernel/
 * load_all_buffers (main.c:249-280): validates the static data, repacks it
 * into the device layout and allocates `members` state slots (1 = the
 * reference's single state; >1 = an ensemble advanced together).  Every
 * stencil index is checked against the patch (SWE_HIP_ERANGE). */
int swe_hip_create(const swe_hip_static *s, int members, unsigned flags,
                   int device, swe_hip_ctx **out);
void swe_hip_destroy(swe_hip_ctx *ctx);

/* FP32 drop-in for the reference built as shipped (RUN_32BIT: the LPSMD
 * arrays and the state are float): an SWE_HIP_FP32 context (members must be
 * 1; flags may add SWE_HIP_EXACT).  Its state goes in and out through
 * swe_hip_set_state32 / swe_hip_get_state32 (host layout SVM_LIN_ID, float)
 * and its RHS hook is swe_hip_eval_rhs32; swe_hip_rk_substep and
 * swe_hip_rk_steps advance it like any context.  The double entry points
 * reject FP32 contexts and the *32 ones FP64 contexts (SWE_HIP_EINVAL). */
int swe_hip_create32(const swe_hip_static32 *s, int members, unsigned flags, int device,
                     swe_hip_ctx **out);
int swe_hip_set_state32(swe_hip_ctx *ctx, int member, const float *H);
int swe_hip_get_state32(swe_hip_ctx *ctx, int member, float *H);
int swe_hip_eval_rhs32(swe_hip_ctx *ctx, int member, const float *K, float *F);

/* Extended creation.  members/flags/device as swe_hip_create; member_block
 * is the number of members one lane group of the ensemble kernel serves
 * (a power of two 2..64 dividing members; 0 = the largest such value; 64
 * with a 31-point stencil = the member-pair kernel, two members per lane);
 * share, when not NULL, is a context of the same mesh whose device copy of
 * the static data is reused (s may then be NULL): the static data stay
 * resident until the last context using them is destroyed.  Out-of-range
 * options fail with SWE_HIP_EINVAL before any device call. */
#define SWE_HIP_KERNEL_AUTO 0   /* single state: RESIDENT up to 256 nodes
                                   per compute unit (with a
                                   Hilbert device order unless another is
                                   asked for), else PAIR with a 31-point
                                   stencil, else QUADW (FP32: see below)     */
#define SWE_HIP_KERNEL_QUAD 1   /* FP32 single state: 4 lanes per node
                                   splitting the stencil slots (fast mode)   */
#define SWE_HIP_KERNEL_LANE 2   /* FP32 single state: 1 lane per node        */
#define SWE_HIP_KERNEL_QUADW 3  /* single state: 4 lanes per node, weights,
                                   ids, geometry loaded once per quad and
                                   shared by DPP                             */
#define SWE_HIP_KERNEL_PAIR 4   /* single state: 2 lanes per node, 16-byte
                                   weight and state loads (31-point stencils;
                                   other widths use QUADW)                   */
#define SWE_HIP_KERNEL_TILE 5   /* single state: the PAIR mapping over tiles
                                   of 128 consecutive device nodes whose
                                   distinct neighbour states are staged in
                                   LDS (31-point stencils, else QUADW; use a
                                   HILBERT node order for compact tiles)     */
#define SWE_HIP_KERNEL_OCT 6    /* FP32 single state: 8 lanes per node
                                   splitting the stencil slots (fast mode)   */
#define SWE_HIP_KERNEL_RESIDENT 7 /* single state (FP64: pair mapping; FP32:
                                   one lane per node, LANE's arithmetic),
                                   31-point stencils, at
                                   most 256 nodes per compute unit (65536
                                   on MI355X): swe_hip_rk_steps runs all its
                                   steps in one cooperative launch, one
                                   workgroup per tile of consecutive device
                                   nodes keeping the tile's weights,
                                   positions, geometry, H and D in registers
                                   and waiting, before every stage, for the
                                   tiles it exchanges rows with; same
                                   arithmetic as PAIR (FP32: LANE), bitwise
                                   in both modes.  A wait
                                   that times out (the state is then
                                   invalid) is reported as SWE_HIP_EDEVICE
                                   by the next swe_hip_synchronize,
                                   swe_hip_get_state, swe_hip_query or
                                   swe_hip_rk_steps after the stream has
                                   drained.  Drop-in substeps and eval_rhs
                                   use PAIR (FP32: LANE).  A wait gives up
                                   after 2^22 polls (seconds); the
                                   environment variable
                                   SWE_HIP_RESIDENT_SPIN_LIMIT, read at
                                   context creation, sets another limit
                                   (1 forces time-outs, for testing the
                                   error path).                             */
/* In SWE_HIP_FP32 contexts: LANE = one lane per node; TILE = one lane per
 * node over LDS tiles of tile_nodes (64, 128 or 256) consecutive device
 * nodes (31-point stencils; use a HILBERT node order); both sum every
 * stencil in order (exact and fast mode).  PAIR, QUAD and OCT = 2, 4 or 8
 * lanes per node, each summing every 2nd, 4th or 8th stencil slot, the
 * partial sums added across the group (fast mode only).  RESIDENT = the
 * resident whole-step kernel with one lane per node (both modes).  AUTO =
 * RESIDENT up to 256 nodes per compute unit, else LANE in exact
 * mode, and in fast mode PAIR up to 262144 nodes (one lane per node would
 * leave the chip under-filled), LANE above.  QUADW is an FP64 mapping. */
#define SWE_HIP_ORDER_HOST 0    /* device node order = the caller's (RCM)    */
#define SWE_HIP_ORDER_MORTON 1  /* device node order = 3-D Morton curve      */
#define SWE_HIP_ORDER_HILBERT 2 /* device node order = 3-D Hilbert curve     */
typedef struct swe_hip_options {
    int members;
    unsigned flags;
    int device;
    int member_block;
    swe_hip_ctx *share;
    int kernel;                 /* SWE_HIP_KERNEL_* (members == 1 only)      */
    int prefetch;               /* stencil-loop schedule: 0 (compiler's),
                                   4 or 8 (ring of slots; member-pair kernel:
                                   ring depth 2 or 4, default 3); member-lane
                                   tiles without SWE_HIP_EXACT: 8 takes the
                                   stencil sums on the FP64 matrix cores
                                   (v_mfma_f64_4x4x4f64)                     */
    int node_order;             /* SWE_HIP_ORDER_* of the device copy (ignored
                                   with share: the mesh's order is used)    */
    int tiled;                  /* 1: LDS-staged ensemble kernel (members
                                   >= 32, member_block 32, 31-point
                                   stencil): persistent member-lane tiles,
                                   loader and compute waves over two LDS
                                   images, 160 KB per workgroup             */
    int slot_block;             /* stencil arrays: 0 slot-major [nnbr][n];
                                   16 node blocks [n/16][nnbr][16] (ignored
                                   with share)                              */
    int tile_nodes;             /* tile shape (0 = default): nodes per tile of
                                   SWE_HIP_KERNEL_TILE (64, 128 or 256; up to
                                   4x as many distinct neighbour rows), or of
                                   the member-lane tiled kernel (1..8)      */
    int tile_rows;              /* distinct rows per member-lane tile (0 =
                                   70; 16..70)                              */
} swe_hip_options;
int swe_hip_create_ex(const swe_hip_static *s, const swe_hip_options *o, swe_hip_ctx **out);

/* Autotuner of the ensemble kernel's member block: times `steps` fused RK
 * steps of a `members`-slot context on the static data of `base` for every
 * admissible member block (powers of two 2..64 dividing members) and
 * returns the fastest in *best_mb with its time per step in
 * *best_ms_per_step.  Timing uses HIP events on the candidates' streams. */
int swe_hip_tune_member_block(swe_hip_ctx *base, int members, unsigned flags, int steps,
                              int *best_mb, double *best_ms_per_step);

/* The same over the LDS-tiled ensemble kernel too (member block 32,
 * member-lane tiles, 31-point stencils, tiled = 1).  Returns the fastest (member block, tiled) pair; tiles follow
 * the node order of base's device copy (create base with
 * SWE_HIP_ORDER_HILBERT for compact tiles).  No counterpart in the
 * reference (its ensemble runs are separate processes). */
int swe_hip_tune_ensemble(swe_hip_ctx *base, int members, unsigned flags, int steps,
                          int *best_mb, int *best_tiled, double *best_ms_per_step);

/* Autotuner over the whole configuration space: times `steps` fused RK
 * steps of every candidate on base's static data and node order and writes
 * the fastest, ready for swe_hip_create_ex (share = base), to *best:
 *   members == 1: the PAIR kernel, the RESIDENT kernel (31-point stencils,
 *     at most 256 nodes per compute unit) and the TILE kernel with tiles of
 *     64, 128 and 256 nodes and each row-staging schedule (prefetch 0:
 *     through registers, 4: direct-to-LDS, 8: direct-to-LDS with the first
 *     weights loaded before the tile barrier);
 *   members > 1: every member block untiled, and member block 32 on
 *     member-lane tiles of tile_nodes 8 or 6 x tile_rows 70 or 56;
 *   SWE_HIP_FP32 contexts (flags must carry SWE_HIP_FP32, members == 1):
 *     LANE, TILE with 64, 128 and 256 nodes, and in fast mode PAIR, QUAD
 *     and OCT (lanes per node), and RESIDENT as above.
 * A candidate that base's mesh or node order cannot take (its create fails
 * with SWE_HIP_EINVAL, e.g. resident tiles of a scattered node order) is
 * passed over.  No counterpart in the reference. */
int swe_hip_tune_ex(swe_hip_ctx *base, int members, unsigned flags, int steps,
                    swe_hip_options *best, double *best_ms_per_step);

/* Per-attempt state upload (main.c:297 + 304-315): H in the host state
 * layout (SVM_LIN_ID, patch_size*4 values) into member slot `member`. */
int swe_hip_set_state(swe_hip_ctx *ctx, int member, const double *H);
/* Read-back (main.c:379-380): member slot -> host state layout. */
int swe_hip_get_state(swe_hip_ctx *ctx, int member, double *H);

/* Exactly RK_substep / RK_substep_ocl for every member: eval_RHS into F,
 * then copy/update_D and eval_K/update_H as separate kernels with the
 * reference's constants (dt/2.0, 2.0, 1.0, dt/6.0).  substep_id 0..3. */
int swe_hip_rk_substep(swe_hip_ctx *ctx, int substep_id);

/* Fast path: nsteps full RK4 steps (4 fused stage kernels per step, the RHS
 * never stored, K double-buffered), replayed from captured HIP graphs.
 * Same arithmetic per node as swe_hip_rk_substep. */
int swe_hip_rk_steps(swe_hip_ctx *ctx, int nsteps);

/* No-allocation variant of swe_hip_rk_steps.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Advances the caller-owned
 * device state H (n_nodes*4 doubles in the context's device layout:
 * [node][u v w h] in the device node order, e.g. a copy of
 * swe_hip_state_device_ptr after swe_hip_set_state) by nsteps fused RK4
 * steps on the context's stream.  Same kernels and arithmetic as
 * swe_hip_rk_steps, launched without a graph; the stage states and the RK
 * sum live in a library-owned 64 MiB device array (one per GPU, held until
 * the process ends) instead of context buffers, so the call makes no device
 * allocation.  Limits: single-state FP64 contexts (members == 1) of
 * at most 699050 nodes (larger meshes fail with SWE_HIP_EINVAL).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * threads, other streams) race. */
int swe_hip_rk_steps_noalloc(swe_hip_ctx *ctx, double *H, int nsteps);

/* Test hook: one eval_RHS_swe (rk4_rbffd_swe.c:189-362) of a host state
 * K into F (host layouts), on member slot `member`. */
int swe_hip_eval_rhs(swe_hip_ctx *ctx, int member, const double *K, double *F);

/* Timing of the drop-in path (wall clock around each synchronised kernel,
 * as the OpenCL path's clFinish-per-launch timing, RK_ocl.c:76). */
int swe_hip_set_timing(swe_hip_ctx *ctx, int on);
int swe_hip_get_timers(swe_hip_ctx *ctx, swe_hip_timers *t, int reset);

int swe_hip_synchronize(swe_hip_ctx *ctx);
/* 1 when all work queued on the context's stream has finished, 0 while
 * some is pending; never blocks. */
int swe_hip_query(swe_hip_ctx *ctx);

/* Ensemble slot management (members > 1): only member slots
 * [0, active) are advanced by the following steps (lane groups wholly
 * beyond `active` skip their work; the states of slots >= active are
 * unspecified afterwards).  Stream-ordered, safe inside captured graphs.
 * move_member copies the state H of slot `from` into slot `to`. */
int swe_hip_set_active(swe_hip_ctx *ctx, int active);
int swe_hip_move_member(swe_hip_ctx *ctx, int from, int to);

/* Introspection for callers that drive the context's stream themselves
 * (benchmarks, the ensemble server): device state pointer of the layout
 * [node][var][member] (double, n_nodes*4*members), the HIP stream as an
 * opaque handle, and the sizes. */
int swe_hip_info(const swe_hip_ctx *ctx, int *n_nodes, int *members,
                 int *nnbr, unsigned *flags);
void *swe_hip_state_device_ptr(swe_hip_ctx *ctx);
void *swe_hip_stream(swe_hip_ctx *ctx);
/* Device bytes held by the context's state buffers, and by the static
 * data it uses (shared between contexts created with `share`). */
long long swe_hip_device_bytes(const swe_hip_ctx *ctx);
long long swe_hip_static_bytes(const swe_hip_ctx *ctx);

const char *swe_hip_strerror(int code);
const char *swe_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
