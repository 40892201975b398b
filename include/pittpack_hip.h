/*
 * pittpack_hip.h -- C ABI of the MI355X (gfx950) PittPack solve path.
 *
 * The entry points mirror the public set-up/solve interface of the reference's
 * PencilDcmp (src/include/pencilDcmp.hpp:441-466) as main.cpp drives it
 * (src/main.cpp:73-118): construct -> assignBoundary -> setBox -> setCoords ->
 * assignRhs -> pittPack -> read the solution -> destroy.  Plain pointers and sizes only.
 *
 * Return values are PittPack's error codes (src/include/params.h:86-114, same numbers and
 * meanings: 0 SUCCESS, 15 ALLOCATION_FAIL, 19 INPUT_ARGS, ...) plus the codes >= 24 below.
 * No entry point calls exit(); every failure is returned.
 *
 * Threading: a handle owns its device buffers and one HIP stream; different handles may be
 * used from different threads; one handle must not be used from two threads at once.
 */
#ifndef PITTPACK_HIP_H
#define PITTPACK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pp_solver *pp_handle;

/* codes beyond the reference's 0..23 */
enum {
    PP_SUCCESS = 0,
    PP_ALLOCATION_FAIL = 15,    /* reference ALLOCATION_FAIL */
    PP_INPUT_ARGS = 19,         /* reference INPUT_ARGS */
    PP_HIP_ERROR = 24,          /* a HIP runtime call failed */
    PP_UNSUPPORTED_BC = 25,     /* boundary-condition combination not available */
    PP_UNSUPPORTED_SIZE = 26,   /* no compiled transform / z-solve for this size */
    PP_INVALID_HANDLE = 27,
    PP_CALL_ORDER = 28,         /* e.g. pittPack before assignBoundary / setCoords / RHS */
    PP_NO_DEVICE = 29,          /* no HIP device visible */
    PP_IO_FAIL = 30,            /* a solution file could not be written */
    PP_NO_HDF5 = 31             /* HDF5 output asked for, but no HDF5 1.10+ library loads */
};

/* solution file formats (PP_OPT_IO_FORMAT, pp_write_solution) */
enum {
    PP_IO_HDF5 = 0, /* the reference's layout: Pxdmf3d<app>.h5 (X, Y, Z, Q) + XDMF index */
    PP_IO_RAW = 1   /* Pxdmf3d<app>.bin (raw little-endian FP64) + XDMF 3DCoRectMesh index */
};

/* option keys for pp_set_option */
enum {
    /* 0: reference row-0 rule for the singular (0,0) column whenever |eig+2| < 1e-10
     *    (src/triDiag.cpp:1753-1776);
     * 1 (default): apply it only when the z operator is singular (z Neumann-Neumann); with a
     *    Dirichlet z face the true system is solved. */
    PP_OPT_ZERO_MODE = 1,
    /* value k > 0: record per-stage HIP events for the last k solves (pp_get_stage_times);
     * 0 switches timing off */
    PP_OPT_TIMING = 2,
    /* z solver: 0 = PCR across 16 lanes then Thomas per lane (cyclic rows);
     *           1 = partitioned Thomas, contiguous rows per lane, cross-lane scans;
     *           2 = multigrid V-cycles (the reference's MultiGrid alternative) to a
     *               max-norm residual of PP_OPT_MG_TOL * max|rhs| (default 1e-16), or until
     *               a V-cycle reduces it by less than 10% (its rounding floor);
     *           3 = full parallel cyclic reduction (TriDiag::pcr, src/triDiag.cpp:1442-1542);
     *           4 = cyclic reduction (TriDiag::crp, src/triDiag.cpp:1285-1437).
     *           FP64 form, non-periodic z; column lengths: 0, 3, 4 up to 1024 rows, 1 up
     *           to 512, 2 up to 1365 (the mixed form runs its own FP32 solve, up to 1024) */
    PP_OPT_ZSOLVER = 3,
    /* 0 (default): FP64 throughout.  1: mixed form -- FP32 transforms and z solves with
     * FP64 twiddles and eigenvalues, FP64 solves of the near-singular columns
     * (|eig+2| < tau), the pinned (0,0) column exactly in FP64, then PP_OPT_REFINE steps of
     * FP64-residual refinement.  Columns up to 1024 rows.  Periodic z: the circulant z
     * solve computes in FP64 on the FP32 field, and the singular column follows the
     * reference's averaged row 0 (shermanMorrisonThomasV1) from FP64 plane sums. */
    PP_OPT_PRECISION = 4,
    PP_OPT_REFINE = 5, /* refinement steps of the mixed form (default 2) */
    PP_OPT_TAU = 6,    /* pp_set_option_double: near-singular threshold (default 0.05) */
    PP_OPT_BATCH_MEMBERS = 7, /* server: most requests per batch (default 256) */
    /* 1 (default) = the y transforms hand the z solve contiguous column tiles through a
     * second field-sized buffer (FP64 form: 16-column FP64 tiles; mixed form: 32-column FP32
     * tiles, where the y length and z length allow); 0 = the z solve works on the field in
     * place */
    PP_OPT_ZLAYOUT = 8,
    PP_OPT_MG_TOL = 9, /* pp_set_option_double: multigrid residual tolerance */
    /* batched form: 1 (default) = the second pp_batch_solve captures the solve's kernel
     * launches into a HIP graph and later solves replay it (one submission per solve);
     * 0 = every solve launches its kernels.  Timing (PP_OPT_TIMING) runs eagerly.
     * Single-solve handle (default 0): 1 = repeated pp_solve_device / pp_solve_device_oop
     * calls on the same device pointers (the flow-solver pattern) run the first call
     * eagerly, capture the second into a graph and replay it from then on; any option, box
     * or boundary change, or a new stream, drops it.  Needs a non-null stream; an in-place
     * mixed-form solve runs eagerly. */
    PP_OPT_GRAPH = 10,
    /* 1 = before the next solve (batched form: at pp_batch_finalize, per line class; the
     * batch's right-hand sides are set afterwards), time the candidate configurations on this
     * shape and keep the fastest; the choice is cached per process and shape.  FP64 form --
     * z: PCR->Thomas on z tiles, the same in place, or partitioned Thomas; each line stage:
     * default or alternative radix plan / block shape.  Mixed form -- FP32 z tiles or the
     * work field in place, and the tiled y shape.  0 (default) = fixed default
     * configuration. */
    PP_OPT_AUTOTUNE = 11,
    /* testing: force every FP64 line stage (and the mixed form's FP32 tiled y stages) to the
     * default (0) or alternative (1) launch shape where it has one; -1 (default) = the
     * autotuner's choice */
    PP_OPT_LINE_VARIANT = 12,
    /* server: how many admitted batches may overtake a queued request that does not fit
     * before nothing behind it is admitted (default 4); bounds the wait of large requests
     * under a stream of small ones */
    PP_OPT_MAX_OVERTAKE = 13,
    /* pp_io file format: PP_IO_HDF5 (default) or PP_IO_RAW */
    PP_OPT_IO_FORMAT = 14,
    /* pp_set_option_double / pp_batch_set_option_double (before finalize), FP64 form with z
     * solver 0 or 1, columns up to 512 rows: columns with |eig+2| below this threshold (not the
     * pinned one) are solved by the reference's Thomas sweeps (thomasLowMem,
     * src/triDiag.cpp:1693-1822, its factors in the reference's order) instead of the PCR
     * steps, which lose up to ten times more digits on nearly singular columns (long Neumann or
     * periodic axes, cells thinner in z).  Default 1e-3 (the columns where the two differ beyond
     * 1e-13); 0 = off. */
    PP_OPT_ZTHOMAS_TAU = 15,
    /* mixed form: each refinement step computes the FP64 residual inside its FP32 x-forward
     * transform (one kernel; the residual is never stored), where the x lines have a compiled
     * plan and ny is a multiple of the kernel's lines per block (8 for 512-point lines); 0 = the
     * residual kernel, then the transform.  Default 1.  Batches: per member by the same rule,
     * set before pp_batch_finalize. */
    PP_OPT_FUSE_RESIDUAL = 16
};

/* PencilDcmp(argc, argv, nx, ny, nz) / make_Poisson (src/pencilDcmp.cpp:232-452,6432-6440).
 * Global grid sizes (one GPU holds the whole grid).  Allocates the device field. */
int pp_create(int nx, int ny, int nz, pp_handle *out);

/* ~PencilDcmp (src/pencilDcmp.cpp:528-658) */
int pp_destroy(pp_handle h);

/* PencilDcmp::assignBoundary(char bc[6]) (src/pencilDcmp.cpp:3578-3634).
 * bc = faces xmin,xmax,ymin,ymax,zmin,zmax, each 'P', 'N' or 'D'.  x and y take PP, NN, DD,
 * and (extension) ND / DN; z takes DD, NN, DN, ND. */
int pp_assign_boundary(pp_handle h, const char *bc);

/* PencilDcmp::setBox(double X[6]) (src/pencilDcmp.cpp:470-480) */
int pp_set_box(pp_handle h, const double *X);

/* PencilDcmp::setCoords(int dir) (src/pencilDcmp.cpp:748-840); dir = 2 as main.cpp uses */
int pp_set_coords(pp_handle h, int dir);

/* PencilDcmp::assignRhs(double *rhs) (src/pencilDcmp.cpp:5591-5647): host array, real,
 * nx*ny*nz, x fastest, unscaled.  Copied to the device field. */
int pp_assign_rhs(pp_handle h, const double *rhs);

/* Same, from a device pointer (nx*ny*nz doubles, x fastest). */
int pp_assign_rhs_device(pp_handle h, const double *d_rhs);

/* PencilDcmp::pittPack() (src/pencilDcmp.cpp:5801-6430): solve in place on the device field.
 * Blocks until the solve has finished on the handle's stream. */
int pp_pittpack(pp_handle h);

/* PencilDcmp::initializeTrigonometric (src/pencilDcmp.cpp:1871-2035, with the reference's
 * SHIFT = 1, EXACT = 0, COEFF = 1, 2, 3): set the handle's right-hand side to the reference's
 * manufactured field on the current box and boundary conditions (after pp_set_coords), in
 * place of pp_assign_rhs.  That field is complex in the reference; part 0 sets its real part,
 * part 1 its imaginary part (two independent real problems).  Cases, as the reference picks
 * them: x, y and z periodic: u = cos / sin(2 pi (x+y+z)); else x or y periodic:
 * u = sin(2 pi z) cos / sin(2 pi (x+y)); else u = prod_d exactValue(omega_d, x_d, tag_d) with
 * omega = (1, 2, 3) pi (imaginary part 0); f = laplacian(u).  Synchronises the stream. */
int pp_init_trigonometric(pp_handle h, int part);

/* PencilDcmp::getError (src/pencilDcmp.cpp:2040-2208): *err = RMS over the grid of the
 * handle's field minus the exact solution of the manufactured case -- after
 * pp_init_trigonometric(part), that part's exact solution; otherwise (right-hand side from
 * pp_assign_rhs, the reference's INITANALYTIC = 0 path) cos(2 pi z), the solution of
 * pp_fill_trigonometric's right-hand side.  Unlike the reference it leaves the field intact.
 * Synchronises the handle's stream. */
int pp_get_error(pp_handle h, double *err);

/* PencilDcmp::fillTrigonometric (src/pencilDcmp.cpp:5685-5765): write the right-hand side
 * -4 pi^2 cos(2 pi z) at the cell centres of the current box into the host array rhs
 * (nx*ny*nz doubles, x fastest), for pp_assign_rhs. */
int pp_fill_trigonometric(pp_handle h, double *rhs);

/* No reference counterpart (P is protected): copy the solution to a host array. */
int pp_get_solution(pp_handle h, double *u);

/* No reference counterpart: copy the solution to a device array. */
int pp_get_solution_device(pp_handle h, double *d_u);

/* PencilDcmp::IO(int app, int dir, int aligndir) (src/pencilDcmp.cpp:1827-1832, writing
 * through Phdf5::writeMultiBlockCellCenter and xdmfMultiBlockCellCenter,
 * src/phdf5.cpp:1299-1808,1809-2237): write the handle's field into <folder> in the format
 * PP_OPT_IO_FORMAT selects (pp_write_solution).  The folder must exist (the reference writes
 * into soln/).  One device holds the whole grid, so the reference's chunk arguments dir /
 * aligndir have no counterpart.  Synchronises the handle's stream.  PP_IO_FAIL when a file
 * cannot be written, PP_NO_HDF5 when HDF5 output is asked for and no HDF5 library loads. */
int pp_io(pp_handle h, const char *folder, int app);

/* Flow-solver form (no reference counterpart): solve in place on a caller-owned device array
 * of nx*ny*nz doubles holding the RHS on entry and the solution on exit, on the handle's
 * stream, without a host synchronisation. */
int pp_solve_device(pp_handle h, double *d_field);

/* Out-of-place device form (no reference counterpart): reads the RHS from d_rhs (kept
 * intact) and leaves the solution in d_u (both nx*ny*nz doubles, x fastest; d_u may equal
 * d_rhs).  Asynchronous on the handle's stream. */
int pp_solve_device_oop(pp_handle h, const double *d_rhs, double *d_u);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The FP64 solve
 * of pp_solve_device_oop without a handle: sizes, boundary conditions bc[6], box X[6] and the
 * PP_OPT_ZERO_MODE rule (0 or 1) are arguments; d_rhs is kept, the solution goes to d_u (may
 * equal d_rhs), asynchronous on `stream` (hipStream_t as void*, null = default stream).  It
 * calls no allocator: its tables live in a fixed 2 MiB device array per GPU, held until the
 * process ends, and the z solve runs on d_u in place (no z-tile buffer).  Limits: the
 * handle's sizes (lines 2..1024, z columns 4..1024, periodic z 2..1024); larger grids get
 * PP_UNSUPPORTED_SIZE.  A call with new sizes, boundary conditions, box or zero-mode rule
 * rebuilds the tables on the host, copies them in and synchronises `stream` (PP_CALL_ORDER
 * under stream capture: make one uncaptured call with the same arguments first); a repeated
 * configuration only enqueues kernels.  Calls that overlap on one GPU race on that array:
 * serialise them. */
int pp_solve_device_oop_noalloc(int nx, int ny, int nz, const char *bc, const double *X, int zero_mode,
                                const double *d_rhs, double *d_u, void *stream);

/* Use the caller's HIP stream (hipStream_t passed as void*) instead of the handle's own. */
int pp_set_stream(pp_handle h, void *stream);

int pp_set_option(pp_handle h, int key, long long value);

int pp_set_option_double(pp_handle h, int key, double value);

/* Mean durations (ms) of the five stages of the FP64 form over the solves recorded since
 * the last call: x forward, y forward, z solve, y inverse, x inverse.  Requires
 * PP_OPT_TIMING.  Synchronises the handle's stream. */
int pp_get_stage_times(pp_handle h, float *ms5);

/* Mean durations (ms) between consecutive kernel launches of the recorded solves (any
 * form), in launch order; *count receives the number of intervals.  Resets the record. */
int pp_get_kernel_times(pp_handle h, float *ms, int max, int *count);

/* Configuration in use: out[0] = 1 when tuned, out[1] = z solver, out[2] = 1 when z tiles
 * are used, out[3..6] = launch shape (0 default, 1 alternative) of the x forward, y forward,
 * y inverse and x inverse stages; with n >= 9, the mixed form's out[7] = 1 when FP32 z tiles
 * are used and out[8] = their y shape (1: half the threads per column).  n >= 7.  No
 * counterpart in the reference. */
int pp_get_tuning(pp_handle h, int *out, int n);

/* Device pointer of the handle's field (nx*ny*nz doubles). */
int pp_get_device_field(pp_handle h, double **d_field);

/* PittPackGetErrorEnum (src/pencilDcmp.cpp:4869-4913) extended with the codes above. */
const char *pp_get_error_enum(int code);

/* Host-side writer behind pp_io (no GPU needed): the field u (nz*ny*nx doubles, x fastest)
 * of the box X[6] as
 *   PP_IO_HDF5: <folder>/Pxdmf3d<app>.h5 holding the reference's four FP64 datasets X, Y, Z
 *     (node coordinates, (nz+1) x (ny+1) x (nx+1) x 1) and Q (nz x ny x nx x 1), each chunked
 *     as one block (src/phdf5.cpp:1380-1420,1700-1760), and <folder>/Pxdmf3d<app>.xmf, an
 *     XDMF 2.0 spatial collection with one 3DSMesh grid whose X_Y_Z geometry and cell
 *     attribute Q are hyperslabs of those datasets (src/phdf5.cpp:1809-2070);
 *   PP_IO_RAW: <folder>/Pxdmf3d<app>.bin (raw little-endian FP64, z slowest) and an XDMF
 *     index describing the same mesh as a 3DCoRectMesh (origin and spacing).
 * HDF5 is loaded at run time (dlopen; $PP_HDF5_LIB first, then libhdf5.so on the loader's
 * path and the image's known locations); PP_NO_HDF5 when none loads. */
int pp_write_solution(const char *folder, int app, int nx, int ny, int nz, const double *X, const double *u,
                      int format);

/* 1 when an HDF5 1.10+ library could be loaded for PP_IO_HDF5 output, else 0 */
int pp_hdf5_available(void);

/* 1 if the transform length n is compiled in, else 0 */
int pp_size_supported(int n);

/* GPU self test of the 16-lane rotation the z solve relies on; 0 when it behaves as assumed */
int pp_selftest(void);

/* No reference counterpart: streaming copy of n doubles between 16-byte aligned device
 * arrays on a HIP stream (mode 0 cached, 1 nontemporal loads and stores, 2 nontemporal stores,
 * 3 nontemporal loads), asynchronous.  The yardstick of the
 * achievable HBM bandwidth (16 bytes per element, as one FP64 stage per grid point). */
int pp_copy_device(const double *d_in, double *d_out, long long n, int mode, void *stream);

/* ---------------------------------------------------------------------------------------
 * Batched form (no reference counterpart; the reference solves one grid per PencilDcmp
 * object, src/pencilDcmp.cpp:5801-6430).  Many independent solves of different sizes,
 * boundary conditions and boxes: their right-hand sides and solutions sit back to back in
 * device arenas and each stage runs as one launch per shape class (line length and
 * transform kind for x and y, rows per lane for z) over all members of the class.  A
 * member's FP64 result is bitwise the one the single-solve handle computes.
 * Order: pp_batch_create, [pp_batch_set_option(PP_OPT_PRECISION / ZERO_MODE),
 * pp_batch_set_option_double(PP_OPT_TAU / PP_OPT_ZTHOMAS_TAU)], pp_batch_add (each member),
 * pp_batch_finalize,
 * fill the right-hand sides (pp_batch_set_rhs / pp_batch_field / pp_batch_fill_uniform),
 * pp_batch_solve (asynchronous on the batch's stream; right-hand sides kept), read
 * (pp_batch_get_solution / pp_batch_solution).
 * ------------------------------------------------------------------------------------- */
typedef struct pp_batch_s *pp_batch;

int pp_batch_create(pp_batch *b);
int pp_batch_destroy(pp_batch b);
/* member with the reference's set-up arguments (sizes, bc[6], box X[6]); *id = its index.
 * Any boundary combination the single-solve handle takes, periodic z included (FP64 members
 * with periodic z keep their field in place instead of z tiles); columns up to 1024 rows. */
int pp_batch_add(pp_batch b, int nx, int ny, int nz, const char *bc, const double *X, int *id);
int pp_batch_finalize(pp_batch b);
int pp_batch_count(pp_batch b, int *members, long long *points);
/* launches per stage: x, y, z (one per shape class) */
int pp_batch_launch_count(pp_batch b, int *per_stage3);
/* device pointers of a member's right-hand side and solution (nx*ny*nz doubles each) */
int pp_batch_field(pp_batch b, int id, double **d_rhs);
int pp_batch_solution(pp_batch b, int id, double **d_u);
int pp_batch_set_rhs(pp_batch b, int id, const double *rhs);
int pp_batch_get_solution(pp_batch b, int id, double *u);
int pp_batch_get_rhs(pp_batch b, int id, double *rhs);
/* every member's right-hand side from the counter generator with seed (seed + id),
 * optionally rounded to FP32 */
int pp_batch_fill_uniform(pp_batch b, unsigned long long seed, int round_f32);
int pp_batch_set_stream(pp_batch b, void *stream);
/* Number of (line class, direction) pairs of an FP64 batch running the alternative launch
 * shape after PP_OPT_AUTOTUNE.  No counterpart in the reference. */
int pp_batch_tuning(pp_batch b, int *alt_shapes);

/* PP_OPT_ZERO_MODE, PP_OPT_PRECISION (before pp_batch_add), PP_OPT_REFINE, PP_OPT_TIMING */
int pp_batch_set_option(pp_batch b, int key, long long value);
int pp_batch_set_option_double(pp_batch b, int key, double value);
int pp_batch_solve(pp_batch b);
/* mean time (ms) between consecutive launch marks of the recorded batch solves
 * (FP64: x fwd, y fwd, z, y inv, x inv) */
int pp_batch_get_kernel_times(pp_batch b, float *ms, int max, int *count);
int pp_batch_synchronize(pp_batch b);

/* ---------------------------------------------------------------------------------------
 * Serving form (no reference counterpart).  A request loop on its own scheduler thread:
 * requests of any supported size, boundary condition and precision are admitted by HBM
 * footprint into a device pool of hbm_bytes, grouped into batches (one per precision,
 * at most PP_OPT_BATCH_MEMBERS requests) and launched on up to `streams` HIP streams, so
 * batches admitted while others run overlap on the device.  Twiddle tables stay resident.
 * A request's result does not depend on the batch it ran in.
 * ------------------------------------------------------------------------------------- */
typedef struct pp_server_s *pp_server;

/* staging_bytes of pinned host memory hold solutions on their way back to callers */
int pp_server_create(long long hbm_bytes, long long staging_bytes, int streams, pp_server *s);
int pp_server_destroy(pp_server s);
/* PP_OPT_REFINE, PP_OPT_BATCH_MEMBERS, PP_OPT_MAX_OVERTAKE */
int pp_server_set_option(pp_server s, int key, long long value);
/* one request: sizes, bc[6], box X[6], precision (0 FP64, 1 mixed), right-hand side rhs
 * (host, nx*ny*nz doubles, x fastest, kept alive until the request is done) or, when rhs is
 * null, the counter generator with `seed`; want_solution = 1 to receive the solution.
 * The request is planned here: unsupported sizes / boundary conditions return their code, and
 * PP_ALLOCATION_FAIL when its exact device footprint exceeds the pool. */
int pp_server_submit(pp_server s, int nx, int ny, int nz, const char *bc, const double *X, int precision,
                     const double *rhs, unsigned long long seed, int want_solution, long long *ticket);
/* blocks until the request is done; fills u (when it was requested and u is not null) and
 * the submit-to-done latency; returns the request's status */
int pp_server_wait(pp_server s, long long ticket, double *u, double *latency_ms);
/* device footprint (bytes of the pool) a request of this shape, boundary and precision takes */
int pp_server_footprint(pp_server s, int nx, int ny, int nz, const char *bc, const double *X, int precision,
                        long long *bytes);
/* the same as pp_server_wait, also returning the launch number of the batch that ran the
 * request */
int pp_server_wait_ex(pp_server s, long long ticket, double *u, double *latency_ms, long long *batch_seq);
int pp_server_stats(pp_server s, long long *completed, long long *points, long long *batches, double *span_ms,
                    int *max_concurrent);

#ifdef __cplusplus
}
#endif

#endif /* PITTPACK_HIP_H */
