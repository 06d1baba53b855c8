/* dragon_gpu.h -- C ABI of the MI355X unsplit Godunov step.
 *
 * A drop-in for the reference's Grid1D/2D/3D step path: the grid state lives
 * in HBM, and boundaries, CFL and the unsplit step run as HIP kernels on one
 * stream per handle.  Plain pointers and sizes only; no exceptions cross this
 * boundary.  Every function returns a DG_* status code.
 *
 * Reference citations are paths under the reference tree (DRAGON/...).
 *
 * Layouts (host and device buffers alike):
 *   cell state   reference ExtendedArray order (Hydro/ExtendedArray/
 *                ExtendedArray.hpp:75,125: last index fastest, ghosts
 *                included), `stride` doubles per cell in PrimitiveState order
 *                rho, p, vx, vy, vz[, Bx, By, Bz] (FluidElement.hpp:34-43).
 *   face B       same order over the (n+1)-sized face index space, 3 doubles
 *                (Bx, By, Bz) per face (Hydro/Grid.hpp:65,107).  2D: .z is
 *                the cell-centred Bz, as in the reference.
 *   SoA (device) field-major: field f of element m at base[f*count + m].
 *
 * Calls on one handle are serialised by the caller; handles are independent.
 */
#ifndef DRAGON_GPU_H
#define DRAGON_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
    DG_OK = 0,
    DG_ERR_UNPHYSICAL = 1,        /* "Unphysical state would be produced at (i,j)"  Hydro/Godunov/Unsplit.cpp:97 */
    DG_ERR_NAN = 2,               /* "NaN state would be produced at (i,j)"          Unsplit.cpp:267 */
    DG_ERR_RIEMANN_INPUT = 3,     /* "Riemann Solution is unphysical because inputs were not physical"  Riemann/Flux.cpp:94 */
    DG_ERR_RIEMANN_SOLUTION = 4,  /* "Riemann Solution to physical inputs is unphysical"                Riemann/Flux.cpp:95 */
    DG_ERR_CFL = 5,               /* "CFL timestep failed: max signal speed is zero"  Hydro/CFL/CFL.cpp:127 */
    DG_ERR_DT_UNDERFLOW = 6,      /* "Timestep has fallen below minimum."  CFL.cpp:230-233 (reference exits) */
    DG_ERR_ARG = 7,
    DG_ERR_HIP = 8,
    DG_ERR_UNSUPPORTED = 9        /* operation not available for this grid, e.g. the split step on MHD (dg_last_error says why) */
};

/* Config.h switches (Config.h:27-110) and the grid geometry (Grid.hpp:46,70,112). */
enum { DG_EXACT = 0, DG_HLL = 1, DG_HLLC = 2, DG_HLLD = 3, DG_HLLE = 4, DG_HLLX = 5 };
enum { DG_MINMOD = 0, DG_MC = 1, DG_VANLEER = 2, DG_SUPERBEE = 3, DG_VANALBADA = 4 };
enum { DG_CT_TOTAL_E = 0, DG_CT_THERMAL = 1, DG_CT_BETA_GATED = 2 };
enum { DG_CFL_MAX = 0, DG_CFL_ADD = 1 };   /* values > 1: power mean, as Config.h:70-77 */
enum { DG_FP64 = 0, DG_MIXED = 1 };        /* precision of reconstruction + Riemann solves */

typedef struct dg_desc {
    int dim;                 /* 1, 2, 3 */
    int nx, ny, nz;          /* interior cells */
    double dx, dy, dz;
    int ghosts;              /* requested; raised as validGhosts (Splitting.cpp:19-27) */
    int mhd;                 /* #define MHD */
    double gamma;            /* Constants.h:18 */
    int solver;              /* RIEMANN_DEFAULT */
    int limiter;             /* MUSCL_DEFAULT_LIMITER */
    int muscl;               /* MUSCL_Hancock */
    int ctu;                 /* CTU */
    int ct_energy;           /* CT_ENERGY_CONSV */
    double ct_beta;          /* CT_Energy_Beta */
    int cfl_mode;            /* CFL_CALCULATION */
    double cfl_coeff;        /* CFL_coeff */
    int verify_fallback;     /* RIEMANN_VERIFY_FALLBACK: -1 = reference default (on iff hydro) */
    int fallback_try_hlle;   /* RIEMANN_FALLBACK_TRY_HLLE */
    double fallback_param;   /* Riemann_ExactFallback_Parameter */
    int hlld_safety;         /* HLLD_PHYSICAL_SAFETY */
    double timestep_tol;     /* Timestep_Tolerance */
    int precision;           /* DG_FP64 (drop-in) or DG_MIXED */
} dg_desc;

/* Boundary kinds (Boundary/Boundary.hpp:160-229) and face bits (:139-149). */
enum { DG_BC_OUTFLOW = 0, DG_BC_PERIODIC = 1, DG_BC_REFLECTIVE = 2, DG_BC_FIXED = 3, DG_BC_IGNORE = 4 };
enum { DG_XN = 1, DG_XP = 2, DG_YN = 4, DG_YP = 8, DG_ZN = 16, DG_ZP = 32 };
typedef struct dg_bc {
    int kind;
    int faces;               /* bit mask; periodic is widened to whole dimensions */
    int corners;             /* corner_ghosts */
    int gated;               /* Outflow::Gated */
    double state[8];         /* Fixed: PrimitiveState rho, p, vx, vy, vz, Bx, By, Bz */
} dg_bc;

typedef struct dg_status {
    int code;                /* DG_* */
    int i, j, k;             /* cell of UNPHYSICAL / NAN */
    char msg[192];           /* the reference's exception text */
} dg_status;

typedef struct dg_stats {
    long steps;              /* accepted steps */
    long failures;           /* rejected attempts (dt halvings) */
    double t_advanced;
} dg_stats;

typedef struct dg_grid dg_grid;

/* new GridND(n..., d..., ghosts) (Splitting.cpp:229-251) on the device.
 * Device code indexes each array in 32 bits: a grid whose face index space,
 * (nx+1+2g)(ny+1+2g)(nz+1+2g), reaches 2^29 elements is refused with
 * DG_ERR_ARG (on one GPU's memory only the largest 1D grids get there). */
int dg_create(const dg_desc* desc, dg_grid** out);
int dg_destroy(dg_grid* g);
/* getGhosts() (Grid.hpp:54,79,121) */
int dg_ghosts(const dg_grid* g);
/* element counts of the cell and face arrays, ghosts included */
long dg_cell_count(const dg_grid* g);
long dg_face_count(const dg_grid* g);

/* Host state in/out: grid[i,j,k] and _B() of the reference (Grid.hpp:52-53,77-83,119-125).
 * stride = doubles per cell record (5 hydro, 8 MHD; more is allowed). B may be NULL. */
int dg_upload(dg_grid* g, const double* w, int stride, const double* B);
int dg_download(dg_grid* g, double* w, int stride, double* B);
/* Device-resident state in/out, SoA (8 cell fields, 3 face fields); no host copy.
 * Ordering: the copies run on the handle's own non-blocking stream (dg_stream),
 * which does not wait for work on any other stream.  The caller must finish
 * writing w_soa / B_soa (dg_upload_device), or reading them from an earlier
 * dg_download_device, before the call: synchronise the producing stream, or
 * enqueue that work on dg_stream(g).  Both calls return after the copies are
 * complete (they synchronise the handle's stream). */
int dg_upload_device(dg_grid* g, const double* w_soa, const double* B_soa);
int dg_download_device(dg_grid* g, double* w_soa, double* B_soa);

/* Grid::boundary = BoundaryList(...) (Boundary/BoundaryList.hpp:22-128): ordered list;
 * uncovered faces get the implicit outflow first. */
int dg_set_boundary(dg_grid* g, const dg_bc* list, int n);
/* boundary.apply(grid) (BoundaryList.hpp:84-93) */
int dg_apply_boundary(dg_grid* g);
/* A boundary list that holds user GhostFill subclasses evaluated on the host
 * (Boundary/GhostFill.hpp:16-30): with a callback set, every boundary pass of
 * the library -- dg_apply_boundary, the passes of dg_advance, and the pass
 * before each directional sweep of dg_split_step (Splitting.cpp:219-299) --
 * calls cb(user) instead of applying the device list.  The callback applies
 * the whole list itself, in the reference's order (the implicit Outflow, then
 * the elements), with dg_apply_boundary_element for built-in elements and
 * dg_download / dg_upload around its own host edits; during a split step
 * these act on the sweep's input state.  It returns 0 or a DG_* code.  NULL
 * clears it. */
int dg_set_boundary_callback(dg_grid* g, int (*cb)(void* user), void* user);
/* One boundary element alone, without the implicit Outflow of the list (the
 * element's apply(), e.g. Outflow.cpp / Periodic.cpp): lets a caller keep the
 * reference's list order when some elements are user GhostFill subclasses
 * evaluated on the host (Boundary/GhostFill.hpp:16-30; BoundaryList.hpp:79-104). */
int dg_apply_boundary_element(dg_grid* g, const dg_bc* bc);
/* CFL::cfl_time(grid) (Hydro/CFL/CFL.cpp:113-162) */
int dg_cfl_dt(dg_grid* g, double* dt, dg_status* st);
/* GridND::unsplit_step(dt) (Unsplit.cpp:28-206, Splitting.cpp:100-115): no commit on failure. */
int dg_step(dg_grid* g, double dt, dg_status* st);
/* dg_step_noalloc: no counterpart in the reference.  For callers that cannot
 * allocate device memory (tight stepping loops, harnesses that forbid
 * hipMalloc): the step of dg_step on caller-owned device buffers, with every
 * scratch array (fluxes, half-step faces, EMFs, the failure word) in one
 * static device array of dg_noalloc_scratch_limit() bytes (64 MiB) per GPU,
 * held until the process ends (enough for 46^3 3D MHD or 446^2 2D MHD).  It
 * calls no hipMalloc / hipFree.
 *   w_in / w_out: SoA cell state, 8 fields of dg_cell_count doubles for the
 *   grid desc describes (ghosts included, as dg_download_device); B_in /
 *   B_out: 3 fields of dg_face_count doubles (2D/3D MHD only, else NULL).
 *   The ghosts of w_in / B_in must be filled; w_out / B_out receive the whole
 *   next state and must not alias the inputs.  stream: hipStream_t or NULL.
 * Grids whose scratch (dg_noalloc_scratch_bytes) exceeds the limit are refused
 * with DG_ERR_ARG before any device call.  Returns after the stream has
 * synchronised; on a failed step the outputs are undefined and st holds the
 * reference's message, as for dg_step.  All calls on one GPU share its array,
 * which keeps the last call's contents: calls that overlap on one GPU (two
 * streams or threads) race.  dg_noalloc_last_error: the message of this
 * thread's last refused call. */
int dg_step_noalloc(const dg_desc* desc, const double* w_in, const double* B_in, double* w_out, double* B_out, double dt,
                    void* stream, dg_status* st);
long dg_noalloc_scratch_bytes(const dg_desc* desc);
long dg_noalloc_scratch_limit(void);
const char* dg_noalloc_last_error(void);
/* Grid::advance_unsplit(dt, check_cfl) (CFL.cpp:213-238); max_steps < 0: no limit. */
int dg_advance(dg_grid* g, double dt, int check_cfl, long max_steps, dg_stats* stats, dg_status* st);
/* Grid2D/3D::split_step(dt) (Hydro/Godunov/Splitting.cpp:118-216): Strang-split
 * directional Godunov sweeps (advanceX/Y/Z, :219-299), the order rotating
 * per accepted step; hydro only (MHD returns DG_ERR_UNSUPPORTED, as
 * MHD/Electric.cpp:17-18 refuses DIMENSION_UNSPLIT off).  1D: dg_step.  No
 * commit on failure.  Grid::advance_split(dt, check_cfl) (Hydro/CFL/CFL.cpp:160-184)
 * is dg_advance_split, the dg_advance loop around dg_split_step. */
int dg_split_step(dg_grid* g, double dt, dg_status* st);
int dg_advance_split(dg_grid* g, double dt, int check_cfl, long max_steps, dg_stats* stats, dg_status* st);
/* Grid2D/3D::initialize_B_fields() (MHD/B_fields.cpp:117-149) */
int dg_init_B(dg_grid* g);
/* Error text of the last failing call on this handle. */
const char* dg_last_error(const dg_grid* g);
/* HIP stream the handle's kernels run on (as hipStream_t). */
void* dg_stream(dg_grid* g);
/* Live timing of the step kernels with HIP events on the handle's stream.
 * enable: 1 start (resets), 0 stop, -1 read only.  step_ms: summed device
 * time from the first to the last kernel of each timed step. */
int dg_timing(dg_grid* g, int enable, double* step_ms, long* steps);
/* Per-kernel-class device time accumulated while timing is on (events after
 * each launch).  id in [0, count): returns the class name ("boundary", "cfl",
 * "pred_sweep_x", ..., "apply"), total ms and launches; DG_ERR_ARG past the
 * last class.  Instrumentation of this library, no reference counterpart. */
int dg_kernel_time(dg_grid* g, int id, const char** name, double* ms, long* launches);
/* Diagnostics: max over interior cells of |div B| * min(dx) / max|B_face| (0 for hydro / 1D). */
int dg_divb(dg_grid* g, double* out);

/* Frames and restarts (the reference's DRAGONHOARD; file I/O in dragon_hoard.h).
 * dg_snapshot: the cells and faces that DRAGONHOARD::writeToFile reads
 * (grid[i,j,k] and _B() over [-ng, n+ng), DRAGONHOARD/HDF5Output.cpp:305-342,
 * 450-491), x fastest in the layout of dragon_hoard.h, copied into the host
 * buffers asynchronously: a device transpose into a staging buffer on the
 * grid's stream, then the device-to-host copy on a copy stream, so later
 * steps overlap the transfer.  cells: 8 x dgh_cell_count doubles; faces: 3 x
 * dgh_face_count (2D/3D MHD; may be NULL otherwise).  Host buffers from
 * dg_host_alloc (pinned) let the copy run without a host-side staging pass;
 * they must stay valid until dg_snapshot_wait returns.  A later snapshot waits
 * for the previous copy.  0 <= ng <= dg_ghosts(g).  dg_snapshot_wait reads no
 * handle state but the copy's event: it may be called from another thread
 * while the caller keeps stepping the grid (a frame writer thread). */
int dg_snapshot(dg_grid* g, int ng, double* cells, double* faces);
int dg_snapshot_wait(dg_grid* g);
/* DRAGONHOARD::loadFromFile's writes into the grid (HDF5Restart.cpp:205-261):
 * the cells and faces of the box [-ng, n+ng) from buffers in the same layout;
 * the rest of the state is kept.  Synchronous. */
int dg_restore(dg_grid* g, int ng, const double* cells, const double* faces);
/* grid.dx/dy/dz = ... as loadFromFile sets them (HDF5Restart.cpp:187-188, 290-292). */
int dg_set_spacing(dg_grid* g, double dx, double dy, double dz);
/* Pinned host memory for snapshot buffers (hipHostMalloc / hipHostFree). */
int dg_host_alloc(long bytes, void** p);
int dg_host_free(void* p);

/* ---------------------------------------------------------------------
 * Batched patches.  Replaces the per-child dispatch of DistGrid<T>::step
 * (Refinement/DistGrid.cpp:291-303, children of DistGrid.hpp:20, each child
 * running Grid2D::unsplit_step on a DRAGONWING thread) for many independent
 * 2D MHD grids (CTU + CT, 3 ghosts, the physics switches of `physics`).  A
 * step is a fixed sequence of launches, each covering every patch; it
 * attempts one step of each patch at the patch's own CFL dt.  The dt halving
 * of Grid::advance (CFL.cpp:151-213) is kept per patch on the device: a
 * patch whose attempt fails keeps its state and retries at half the dt on
 * the next step (steps counts committed steps, fails rejected attempts).
 * Unlike DistGrid children, patches do not share a dt or halos.
 * Boundaries: bc_kind DG_BC_OUTFLOW or DG_BC_PERIODIC on every face of every
 * patch, corner ghosts on.  Patch sides must be >= 4 cells.
 * --------------------------------------------------------------------- */
typedef struct dg_batch dg_batch;
int dg_batch_create(const dg_desc* physics, int npatch, const int* nx, const int* ny, const double* dx, const double* dy,
                    int bc_kind, dg_batch** out);
int dg_batch_destroy(dg_batch* b);
/* Interior zones summed over patches; ghost width (3). */
long dg_batch_zones(const dg_batch* b);
int dg_batch_ghosts(const dg_batch* b);
/* Patch state in the single-grid record layout (dg_upload): cells incl.
 * ghosts, stride 8, and face B (3 per face).  Upload makes it current. */
int dg_batch_upload(dg_batch* b, int patch, const double* w, int stride, const double* B);
int dg_batch_download(dg_batch* b, int patch, double* w, int stride, double* B);
/* initialize_B_fields (MHD/B_fields.cpp:117-149) on every patch. */
int dg_batch_init_B(dg_batch* b);
/* nsteps steps of every patch, asynchronous on the batch stream; use_graph:
 * the nsteps launches are captured once into a HIP graph and replayed.
 * device_ms (optional): wait and return the device time of the nsteps. */
int dg_batch_run(dg_batch* b, int nsteps, int use_graph, double* device_ms);
/* Wait; DG_OK, or the status of the first patch whose last step could not
 * complete (st->i = patch index; the patch keeps its previous state). */
int dg_batch_sync(dg_batch* b, dg_status* st);
int dg_batch_patch_stats(dg_batch* b, int patch, double* dt_last, long* steps, long* fails, int* status);
/* Size classes: largest side < 64, < 128, < 192, >= 192.  Two tile shapes
 * per class: the number of x faces one thread walks in the x sweep (tiles),
 * and the y sweep's segment width, 64, 32 or 16 lanes (one, two or four
 * segments of 63, 31 or 15 faces per wavefront; narrow patches waste fewer
 * lanes with short segments), or 0: the patch's columns streamed through its
 * wavefronts, 63 cells each, with no segment boundary inside a column.
 * autotune times each candidate on the class's
 * patches (steps not committed), the x tile first, then the y width, and
 * keeps the fastest; class_ms (optional) gets the best time per launch.
 * Results do not depend on either shape. */
int dg_batch_set_tiles(dg_batch* b, const int* chunk, int n);
int dg_batch_get_tiles(const dg_batch* b, int* chunk, int n);
int dg_batch_set_ysweep(dg_batch* b, const int* width, int n);
int dg_batch_get_ysweep(const dg_batch* b, int* width, int n);
int dg_batch_autotune(dg_batch* b, int reps, int* chunk_out, double* class_ms);
const char* dg_batch_last_error(const dg_batch* b);
void* dg_batch_stream(dg_batch* b);

#ifdef __cplusplus
}
#endif
#endif
