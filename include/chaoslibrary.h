/*
 * chaoslibrary.h -- C-ABI of the MI355X (gfx950) ChaosLibrary hot path.
 *
 * The reference exposes a C++ host API only (no FFI).  Each file-writing
 * entry point below replaces one of its C++ functions with the same argument
 * list, in the same order, plus a leading `const chaos_options*` (NULL =
 * defaults) and a `const char*` path instead of std::string:
 *
 *   chaos_bifurcation1D        <- Bifurcation::bifurcation1D
 *                                 (Bifurcation/include/bifurcationHOST.h:16-30)
 *   chaos_bifurcation2D        <- Bifurcation::bifurcation2D
 *                                 (Bifurcation/include/bifurcationHOST.h:35-50)
 *   chaos_LLE2D                <- LLE::LLE2D (LLE/include/LLEHost.h:25-41)
 *   chaos_basinsOfAttraction_2 <- Basins::basinsOfAttraction_2
 *                                 (Basins/include/basinsHOST.h:15-30)
 *
 * The compute entry points (chaos_bif1d, chaos_bif2d, chaos_lle2d,
 * chaos_basins2d, chaos_integrate) return the same values into caller
 * buffers instead of files; they are what the file writers call.
 *
 * Conventions: every function returns 0 (CHAOS_OK) or a negative chaos_status;
 * chaos_last_error() gives a thread-local message for the last failure.
 * The reference instead prints and exit()s (debri/cudaMacros.cu:3-10).
 * Inputs are borrowed for the duration of the call; all entry points are
 * reentrant (no module-global device symbols) and never reset the device.
 * Plain pointers and sizes only; hip types appear as void*.
 */
#ifndef CHAOSLIBRARY_H
#define CHAOSLIBRARY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHAOS_API_VERSION 1

typedef enum chaos_status {
    CHAOS_OK            = 0,
    CHAOS_E_INVALID     = -1,  /* bad argument                                  */
    CHAOS_E_HIP         = -2,  /* HIP runtime error                             */
    CHAOS_E_NOMEM       = -3,  /* host or device allocation failed             */
    CHAOS_E_IO          = -4,  /* output file could not be opened / written    */
    CHAOS_E_CAPACITY    = -5,  /* caller buffer too small (size reported back) */
    CHAOS_E_NODEVICE    = -6,  /* no HIP device                                */
    CHAOS_E_UNSUPPORTED = -7,  /* combination not implemented                  */
    CHAOS_E_BUSY        = -8,  /* server queue full / not admitted             */
    CHAOS_E_NOTFOUND    = -9   /* unknown job id                               */
} chaos_status;

/* Systems: the five reference systems of systems/systems.cuh (one chosen per
 * reference build by #define USE_*_MODEL, systems/systems.cuh:9) and the two
 * RK4 systems of the benchmark configurations, selected at run time. */
typedef enum chaos_system {
    CHAOS_SYS_CHAMELEON   = 0, /* NX 3, NA 7  systems/systems.cuh:11-39   */
    CHAOS_SYS_ROSSLER_CD  = 1, /* NX 3, NA 4  systems/systems.cuh:41-63   */
    CHAOS_SYS_BASINS      = 2, /* NX 3, NA 5  systems/systems.cuh:67-83   */
    CHAOS_SYS_BASINS_2    = 3, /* NX 2, NA 2  systems/systems.cuh:85-101  */
    CHAOS_SYS_DAMIR       = 4, /* NX 3, NA 2  systems/systems.cuh:103-178 */
    CHAOS_SYS_ROSSLER_RK4 = 5, /* NX 3, NA 3  (a, b, c)                   */
    CHAOS_SYS_LORENZ_RK4  = 6, /* NX 3, NA 3  (sigma, rho, beta)          */
    CHAOS_SYS_COUNT       = 7
} chaos_system;

/* REFERENCE replays the reference's semantics (int(T/h) horizons, FP32
 * divergence tests, quirky per-point dbscan, single-linkage basins);
 * CONFIG uses the benchmark-config semantics (FP64 tests, llround horizons,
 * 1-D DBSCAN peak counts, z-scored 3-D DBSCAN basins). */
typedef enum chaos_semantics {
    CHAOS_SEM_REFERENCE = 0,
    CHAOS_SEM_CONFIG    = 1
} chaos_semantics;

typedef struct chaos_stats {
    double kernel_ms;       /* device time of the integration kernel(s)      */
    double analysis_ms;     /* device time of peak/cluster stages           */
    double total_ms;        /* host wall time of the call                   */
    double rk4_steps;       /* trajectory steps integrated (all points)     */
    long long peaks;        /* peaks stored (bifurcation)                   */
    int launches;           /* kernel launches issued                       */
    int reserved;
} chaos_stats;

typedef struct chaos_options {
    int system;             /* chaos_system; default CHAOS_SYS_CHAMELEON     */
    int semantics;          /* chaos_semantics; default REFERENCE            */
    uint64_t seed;          /* LLE perturbation seed; default 12345          */
    int device;             /* HIP device ordinal; default 0                 */
    int device_buffers;     /* compute outputs are device pointers           */
    void* stream;           /* hipStream_t to use; NULL = library stream     */
    int min_pts;            /* CONFIG basins DBSCAN minPts; default 20       */
    int block;              /* threads per block (0 = tuned default)         */
    chaos_stats* stats;     /* optional timing/work report                   */
    int output;             /* map files (2-D outputs): CHAOS_OUT_CSV (default,
                               0 counts as CSV) and/or CHAOS_OUT_BINARY, a
                               sidecar <path>.bin that chaos_map_bin_to_csv
                               turns into the same text later                */
    int reserved;
} chaos_options;

enum { CHAOS_OUT_CSV = 1, CHAOS_OUT_BINARY = 2 };

/* Writes the CSV text of a map sidecar (<path>.bin from CHAOS_OUT_BINARY):
 * byte-identical to the CSV the entry point writes itself. */
int chaos_map_bin_to_csv(const char* bin_path, const char* csv_path);

/* Fill opt with the process defaults (CHAMELEON, REFERENCE, seed 12345,
 * device 0, minPts 20).  chaos_set_default_options replaces the defaults used
 * when an entry point receives opt == NULL (the C++ API always does); it is
 * the run-time counterpart of the reference's compile-time system choice
 * (#define USE_*_MODEL, systems/systems.cuh:9). */
void        chaos_options_init(chaos_options* opt);
int         chaos_set_default_options(const chaos_options* opt);
/* Where calls made on this thread with opt == NULL (the C++ API) report their
 * chaos_stats; NULL (the default) turns the report off.  The process
 * defaults never carry a stats pointer. */
int         chaos_set_thread_stats(chaos_stats* stats);
int         chaos_api_version(void);
const char* chaos_last_error(void);
int         chaos_system_dims(int system, int* nx, int* na);
int         chaos_device_count(int* count);

/* Device memory helpers (plain pointers) for callers without a framework. */
int chaos_device_malloc(void** ptr, size_t bytes, int device);
int chaos_device_free(void* ptr);
int chaos_memcpy_h2d(void* dst, const void* src, size_t bytes);
int chaos_memcpy_d2h(void* dst, const void* src, size_t bytes);
int chaos_device_synchronize(int device);
/* Device-to-device streaming copy (bytes a multiple of 16); also the
 * known-size reference kernel for HBM traffic counters. */
int chaos_device_copy(void* dst, const void* src, size_t bytes, int device);

/* Advance n trajectories nsteps each: states[n][NX] in/out, params[n][NA].
 * The per-step map is the selected system's calcDiscreteModel
 * (systems/systems.cuh) or RK4 step. */
int chaos_integrate(const chaos_options* opt, int n, const double* params,
                    double* states, long long nsteps, double h);

/* Peak finder over caller rows (Bifurcation/src/bifurcationCUDA.cu:469-514,
 * peakFinder): data[rows][len]; counts[r] = peaks kept after dropping the
 * first, peaks[r*len+k] their values, intervals[r*len+k] = (t[k+1]-t[k])*h. */
int chaos_peak_finder(const chaos_options* opt, const double* data, int rows,
                      int len, double h, int* counts, double* peaks,
                      double* intervals);

/* ---------------- LLE (LLE/src/LLECUDA.cu:205-428) ---------------------- */
/* out[j*nPts + i] = LLE(A_i, B_j), A = linspace(ranges[0], ranges[1], nPts)
 * on values[indicesOfMutVars[0]], B likewise on ranges[2..3] /
 * indicesOfMutVars[1].  writableVar and maxValue are ignored, as in the
 * reference. */
int chaos_lle2d(const chaos_options* opt, double tMax, double NT, int nPts,
                double h, double eps, const double* initialConditions,
                int amountOfInitialConditions, const double* ranges,
                const int* indicesOfMutVars, int writableVar, double maxValue,
                double transientTime, const double* values, int amountOfValues,
                double* out);
int chaos_LLE2D(const chaos_options* opt, double tMax, double NT, int nPts,
                double h, double eps, const double* initialConditions,
                int amountOfInitialConditions, const double* ranges,
                const int* indicesOfMutVars, int writableVar, double maxValue,
                double transientTime, const double* values, int amountOfValues,
                const char* OUT_FILE_PATH);

/* ---------------- Bifurcation (Bifurcation/src/bifurcationCUDA.cu) -------- */
/* 1-D: counts[p] = kept peaks of point p (>= 0), or -1 for a fixed point
 * (first[p] then holds its first recorded sample).  peaks/intervals receive
 * the kept peaks of all points concatenated in point order; capacity is
 * their length.  *total receives the number needed; CHAOS_E_CAPACITY if it
 * exceeds capacity (counts/first/total are still valid).  Host pointers. */
int chaos_bif1d(const chaos_options* opt, double tMax, int nPts, double h,
                int amountOfInitialConditions, const double* initialConditions,
                const double* ranges, const int* indicesOfMutVars,
                int writableVar, double maxValue, double transientTime,
                const double* values, int amountOfValues, int preScaller,
                int* counts, double* first, double* peaks, double* intervals,
                long long capacity, long long* total);
int chaos_bifurcation1D(const chaos_options* opt, double tMax, int nPts,
                        double h, int amountOfInitialConditions,
                        const double* initialConditions, const double* ranges,
                        const int* indicesOfMutVars, int writableVar,
                        double maxValue, double transientTime,
                        const double* values, int amountOfValues,
                        int preScaller, const char* OUT_FILE_PATH);

/* 2-D: out[j*nPts + i] (i over ranges[0..1], j over ranges[2..3]) is -1 for
 * a fixed point, 0 for a divergent or peak-less point, else the cluster
 * count (REFERENCE: the reference's per-point dbscan on intervals; CONFIG:
 * number of 1-D DBSCAN(eps, minPts=1) clusters of the peak values).  A grid
 * whose peak store does not fit the device at once runs in ranges of points
 * that do, with the same result; the environment variable
 * CHAOS_PEAK_STORE_BYTES, read at each call, caps the store's bytes. */
int chaos_bif2d(const chaos_options* opt, double tMax, int nPts, double h,
                int amountOfInitialConditions, const double* initialConditions,
                const double* ranges, const int* indicesOfMutVars,
                int writableVar, double maxValue, double transientTime,
                const double* values, int amountOfValues, int preScaller,
                double eps, int* out);
int chaos_bifurcation2D(const chaos_options* opt, double tMax, int nPts,
                        double h, int amountOfInitialConditions,
                        const double* initialConditions, const double* ranges,
                        const int* indicesOfMutVars, int writableVar,
                        double maxValue, double transientTime,
                        const double* values, int amountOfValues,
                        int preScaller, double eps, const char* OUT_FILE_PATH);

/* ---------------- Basins (Basins/src/basinsCUDA.cu:442-813) ------------- */
/* labels[p], p = j*nPts + i over the IC grid (X[indicesOfMutVars[0]] over
 * ranges[0..1], X[indicesOfMutVars[1]] over ranges[2..3]); flags[p] in
 * {1,0,-1}; features f[k*nPts*nPts + p] (REFERENCE k<2: average peak,
 * average interval with the reference's sentinels; CONFIG k<3: mean X[0],
 * mean X[NX-1], std of the kept peaks -- before the z-score that feeds the
 * clustering).  features/flags may be NULL. */
int chaos_basins2d(const chaos_options* opt, double tMax, int nPts, double h,
                   int amountOfInitialConditions,
                   const double* initialConditions, const double* ranges,
                   const int* indicesOfMutVars, int writableVar,
                   double maxValue, double transientTime, const double* values,
                   int amountOfValues, int preScaller, double eps, int* labels,
                   double* features, int* flags);
int chaos_basinsOfAttraction_2(const chaos_options* opt, double tMax, int nPts,
                               double h, int amountOfInitialConditions,
                               const double* initialConditions,
                               const double* ranges,
                               const int* indicesOfMutVars, int writableVar,
                               double maxValue, double transientTime,
                               const double* values, int amountOfValues,
                               int preScaller, double eps,
                               const char* OUT_FILE_PATH);

/* The clustering stage of Basins on its own, over caller points.
 * Replaces CUDA_dbscan (Basins/src/basinsCUDA.cu:866-923) when sign_rule is
 * set: dim 2, min_pts 1, include[i] == -1 marks a fixed point, and clusters
 * holding one are numbered -1..-K by their smallest fixed-point index, the
 * others 1..K by their smallest index.  With sign_rule clear it is the
 * z-scored 3-D DBSCAN of CONFIG basins: dim 2 or 3, core = at least min_pts
 * points (itself included) within sqrt(sum d^2) <= eps; clusters numbered
 * 1..K by smallest core index; a border point takes the cluster of its
 * smallest-index core neighbour; noise and points with include[i] == 0 get 0.
 * coords: dim arrays of n doubles (x, y[, z]); include may be NULL (all in).
 * With opt->device_buffers the coordinates, include and labels are device
 * pointers.  *clusters (may be NULL) receives K of the positive clusters. */
int chaos_dbscan(const chaos_options* opt, const double* x, const double* y,
                 const double* z, long long n, int dim, double eps,
                 int min_pts, int sign_rule, const int* include, int* labels,
                 int* clusters);

/* No counterpart in the reference.  chaos_dbscan for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops): the same
 * kernels and results, with no hipMalloc, hipFree or other allocation -- the
 * clustering scratch is a static device array of the library (about 48 MB
 * per GPU, held until the process ends).  x, y, z, include and labels are
 * device pointers whatever opt->device_buffers says; the call runs on
 * opt->stream (NULL: the null stream) and fills no opt->stats.  At most
 * CHAOS_DBSCAN_NOALLOC_MAX_POINTS points, and the points may occupy at most
 * 2^17 distinct grid cells (cells of side eps/sqrt(dim)); beyond either it
 * returns CHAOS_E_CAPACITY.  Calls on one GPU share its scratch: calls that
 * overlap on one GPU race. */
#define CHAOS_DBSCAN_NOALLOC_MAX_POINTS (1LL << 20)
int chaos_dbscan_noalloc(const chaos_options* opt, const double* x,
                         const double* y, const double* z, long long n,
                         int dim, double eps, int min_pts, int sign_rule,
                         const int* include, int* labels, int* clusters);

/* ---------------- Legacy entry points (debri/, namespace old_library) ---- */
/* The legacy library sweeps the same integrator with its own LLE kernels
 * (debri/cudaLibrary.cu:1897-2137), the Lyapunov spectrum
 * (debri/cudaLibrary.cu:2189-2331) and a trajectory recorder
 * (debri/hostLibrary.cu:5094-5331).  Windows int(tMax/NT) of int(NT/h)
 * steps; every trajectory stops for the rest of a window after a step that
 * leaves X[0] NaN/inf or fabsf(X[0]) > maxValue (maxValue != 0), as the
 * legacy loop does (writableVar is not used by these kernels).  Directions:
 * seeded unit vectors (chaos_options.seed) instead of curand / sinf.
 *
 *   chaos_LLE1D   <- old_library::LLE1D   (debri/hostLibrary.cuh:241-256)
 *   chaos_LLE1DIC <- old_library::LLE1DIC (debri/hostLibrary.cuh:263-278)
 *   chaos_LLE2DIC <- old_library::LLE2DIC (debri/hostLibrary.cuh:339-354)
 *   chaos_LS1D    <- old_library::LS1D    (debri/hostLibrary.cuh:377-392)
 *   chaos_LS2D    <- old_library::LS2D    (debri/hostLibrary.cuh:416-431)
 *   chaos_TimeDomainCalculation <- old_library::TimeDomainCalculation
 *                                          (debri/hostLibrary.cuh:69-83)
 *
 * chaos_legacy_lle: out[p], p = k0 + nPts*k1 over dims 1 or 2; sweep_ic 0
 * sweeps values[indicesOfMutVars[d]] (LLEKernelCUDA form), 1 sweeps
 * initialConditions[indicesOfMutVars[d]] (LLEKernelICCUDA form).
 * chaos_legacy_ls: out[p*NX + k], exponents of the parameter sweep. */
int chaos_legacy_lle(const chaos_options* opt, double tMax, double NT, int nPts,
                     double h, double eps, const double* initialConditions,
                     int amountOfInitialConditions, const double* ranges,
                     const int* indicesOfMutVars, int writableVar,
                     double maxValue, double transientTime, const double* values,
                     int amountOfValues, int dims, int sweep_ic, double* out);
int chaos_legacy_ls(const chaos_options* opt, double tMax, double NT, int nPts,
                    double h, double eps, const double* initialConditions,
                    int amountOfInitialConditions, const double* ranges,
                    const int* indicesOfMutVars, int writableVar,
                    double maxValue, double transientTime, const double* values,
                    int amountOfValues, int dims, double* out);
int chaos_LLE1D(const chaos_options* opt, double tMax, double NT, int nPts,
                double h, double eps, const double* initialConditions,
                int amountOfInitialConditions, const double* ranges,
                const int* indicesOfMutVars, int writableVar, double maxValue,
                double transientTime, const double* values, int amountOfValues,
                const char* OUT_FILE_PATH);
int chaos_LLE1DIC(const chaos_options* opt, double tMax, double NT, int nPts,
                  double h, double eps, const double* initialConditions,
                  int amountOfInitialConditions, const double* ranges,
                  const int* indicesOfMutVars, int writableVar, double maxValue,
                  double transientTime, const double* values, int amountOfValues,
                  const char* OUT_FILE_PATH);
int chaos_LLE2DIC(const chaos_options* opt, double tMax, double NT, int nPts,
                  double h, double eps, const double* initialConditions,
                  int amountOfInitialConditions, const double* ranges,
                  const int* indicesOfMutVars, int writableVar, double maxValue,
                  double transientTime, const double* values, int amountOfValues,
                  const char* OUT_FILE_PATH);
int chaos_LS1D(const chaos_options* opt, double tMax, double NT, int nPts,
               double h, double eps, const double* initialConditions,
               int amountOfInitialConditions, const double* ranges,
               const int* indicesOfMutVars, int writableVar, double maxValue,
               double transientTime, const double* values, int amountOfValues,
               const char* OUT_FILE_PATH);
int chaos_LS2D(const chaos_options* opt, double tMax, double NT, int nPts,
               double h, double eps, const double* initialConditions,
               int amountOfInitialConditions, const double* ranges,
               const int* indicesOfMutVars, int writableVar, double maxValue,
               double transientTime, const double* values, int amountOfValues,
               const char* OUT_FILE_PATH);
/* Bifurcation over initial conditions (old_library::bifurcation1DIC,
 * debri/hostLibrary.cuh:109-123; bifurcation2DIC, :219-234).  The swept
 * entries are initialConditions[indicesOfMutVars[d]]; the transient
 * (int(transientTime/h) single steps) tests X[0] for NaN/inf only and a
 * fixed point at its end is still recorded; the record runs preScaller
 * steps per sample with the legacy checks (isnan/isinf, fabsf > maxValue).
 * chaos_legacy_bif1d_ic: counts[p] = kept peaks of flag-1 points, else the
 * flag (0 or -1); peaks concatenated (capacity / *total as chaos_bif1d).
 * chaos_legacy_bif2d_ic: out[p] as chaos_bif2d (REFERENCE: the per-point
 * dbscan, debri/cudaLibrary.cu:1593-1666). */
int chaos_legacy_bif1d_ic(const chaos_options* opt, double tMax, int nPts,
                          double h, int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges,
                          const int* indicesOfMutVars, int writableVar,
                          double maxValue, double transientTime,
                          const double* values, int amountOfValues,
                          int preScaller, int* counts, double* peaks,
                          long long capacity, long long* total);
int chaos_legacy_bif2d_ic(const chaos_options* opt, double tMax, int nPts,
                          double h, int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges,
                          const int* indicesOfMutVars, int writableVar,
                          double maxValue, double transientTime,
                          const double* values, int amountOfValues,
                          int preScaller, double eps, int* out);
int chaos_bifurcation1DIC(const chaos_options* opt, double tMax, int nPts,
                          double h, int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges,
                          const int* indicesOfMutVars, int writableVar,
                          double maxValue, double transientTime,
                          const double* values, int amountOfValues,
                          int preScaller, const char* OUT_FILE_PATH);
int chaos_bifurcation2DIC(const chaos_options* opt, double tMax, int nPts,
                          double h, int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges,
                          const int* indicesOfMutVars, int writableVar,
                          double maxValue, double transientTime,
                          const double* values, int amountOfValues,
                          int preScaller, double eps, const char* OUT_FILE_PATH);
/* Recorded samples X[writableVar] of a 1-D parameter sweep: out[p*B + t],
 * B = int(tMax/h/preScaller) (*samples); the transient runs
 * int(transientTime/h) iterations of preScaller steps; flags[p] in
 * {1, 0, -1}; only flag-1 points are recorded (others, and samples after a
 * divergence, are 0).  out == NULL only reports *samples. */
int chaos_time_domain(const chaos_options* opt, double tMax, int nPts, double h,
                      int amountOfInitialConditions,
                      const double* initialConditions, const double* ranges,
                      const int* indicesOfMutVars, int writableVar,
                      double maxValue, double transientTime,
                      const double* values, int amountOfValues, int preScaller,
                      double* out, int* flags, long long capacity,
                      long long* samples);
int chaos_TimeDomainCalculation(const chaos_options* opt, double tMax, int nPts,
                                double h, int amountOfInitialConditions,
                                const double* initialConditions,
                                const double* ranges,
                                const int* indicesOfMutVars, int writableVar,
                                double maxValue, double transientTime,
                                const double* values, int amountOfValues,
                                int preScaller, const char* OUT_FILE_PATH);

/* ---------------- Batched engine and job server ------------------------- */
/* No counterpart in the reference (one synchronous call per sweep, one
 * system per build).  Jobs of any of the four kinds run together: every
 * round launches one persistent kernel per job family (program x system x
 * recorded variable), each on its own stream, whose wavefronts pull (job,
 * 64-point group, tick range) tiles from the family's device queue; lane
 * state stays in HBM between rounds; a round's launch sequence is replayed
 * from a hipGraph captured per set of families.  Jobs are admitted FIFO as
 * job slots, the HBM budget and the tile queue allow (continuous batching),
 * and finished by two worker threads while later rounds integrate.  The
 * engine runs the RK4 systems (CHAOS_SYS_ROSSLER_RK4, CHAOS_SYS_LORENZ_RK4);
 * results are bit-identical to the one-shot entry points for the same job. */
typedef enum chaos_job_kind {
    CHAOS_JOB_LLE2D    = 0,   /* = chaos_lle2d      */
    CHAOS_JOB_BIF1D    = 1,   /* = chaos_bif1d      */
    CHAOS_JOB_BIF2D    = 2,   /* = chaos_bif2d      */
    CHAOS_JOB_BASINS2D = 3    /* = chaos_basins2d   */
} chaos_job_kind;

typedef struct chaos_job {
    int kind;                      /* chaos_job_kind                          */
    int system;                    /* chaos_system (RK4 systems only)         */
    int semantics;                 /* chaos_semantics                         */
    int nPts;
    double tMax, NT, h, eps, maxValue, transientTime;
    int preScaller, writableVar, min_pts, amountOfInitialConditions;
    uint64_t seed;
    double initialConditions[4];
    double values[8];
    int amountOfValues;
    int indicesOfMutVars[2];
    int reserved;
    double ranges[4];
} chaos_job;

typedef struct chaos_engine_params {
    int block;              /* threads per persistent workgroup (64|128|256)  */
    int waves_per_simd;     /* persistent waves per SIMD (1..8); a family with
                               fewer claims runs one wave per claim          */
    int steps_per_launch;   /* RK4 steps a tile advances per launch           */
    int max_jobs;           /* active job slots                               */
    int max_tiles;          /* tile queue capacity per launch                 */
    int use_graph;          /* replay the launch sequence from a hipGraph     */
    double hbm_fraction;    /* share of free HBM admitted jobs may occupy     */
    int points_per_tile;    /* points one wave takes per tile: 64, 128 or 256 */
    int reserved;
} chaos_engine_params;

typedef struct chaos_job_info {
    int status;             /* 0 pending, 1 integrating, 3 finishing (post
                               stages on one of the server's worker threads),
                               2 done, -1 failed                             */
    int kind;
    double t_submit_ms, t_admit_ms, t_done_ms;   /* server clock              */
    double rk4_steps;
    long long footprint_bytes;
    long long n_points;
    long long result_len;   /* 1-D: kept peaks; others: n_points             */
    int reran;              /* peak store overflowed: rerun one-shot          */
    int reserved;
} chaos_job_info;

typedef struct chaos_server_stats {
    long long launches;     /* engine launches (graph replays included)       */
    long long graph_replays;
    long long tiles;
    double busy_ms;         /* device time of the engine launches             */
    double finalize_ms;     /* host wall time in per-job finishing stages     */
    double rk4_steps;       /* steps completed                                */
    int active_jobs, pending_jobs;
    long long hbm_in_use, hbm_budget;
    int finishing_jobs;     /* handed to the worker threads, not yet done     */
    int reserved;
} chaos_server_stats;

/* Defaults: block 256, 4 waves per SIMD, 40000 steps per launch, 64 points
 * per tile, graph replay, 256 job slots, 65536 tiles, 60 % of free HBM (the
 * autotuner's choice on an MI355X; chaos_autotune measures others). */
int chaos_engine_params_default(chaos_engine_params* p);
int chaos_server_create(int device, const chaos_engine_params* p, void** server);
int chaos_server_destroy(void* server);
int chaos_server_submit(void* server, const chaos_job* job, long long* job_id);
/* One scheduling round: admit, launch every active job's next tiles and hand
 * jobs whose integration ended to the server's worker threads, which run
 * their post stages on their own streams while later rounds integrate.
 * *completed receives the number of jobs that became done (or failed) since
 * the previous call.  drain() returns once every job is done or failed. */
int chaos_server_step(void* server, int* completed);
int chaos_server_drain(void* server);
int chaos_server_job_info(void* server, long long job_id, chaos_job_info* info);
int chaos_server_stats_get(void* server, chaos_server_stats* stats);
/* Results (host buffers, same layouts as the one-shot entry points). */
int chaos_server_result_lle(void* server, long long job_id, double* out);
int chaos_server_result_bif2d(void* server, long long job_id, int* out);
int chaos_server_result_bif1d(void* server, long long job_id, int* counts,
                              double* first, double* peaks, double* intervals,
                              long long capacity, long long* total);
int chaos_server_result_basins(void* server, long long job_id, int* labels,
                               double* features, int* flags);
/* Writes a finished job's results to OUT_FILE_PATH in the files and formats
 * of the matching file-writing entry point (chaos_LLE2D, chaos_bifurcation1D,
 * chaos_bifurcation2D, chaos_basinsOfAttraction_2 -- the reference's
 * LLE/src/LLECUDA.cu:383-398, Bifurcation/src/bifurcationCUDA.cu:220-233,
 * 746-763 and Basins/src/basinsCUDA.cu:557-590, 708-792 writers); output is
 * CHAOS_OUT_* as in chaos_options.output (0 = CSV).  The bytes equal the
 * one-shot entry point's for the same job. */
int chaos_server_write_result(void* server, long long job_id,
                              const char* OUT_FILE_PATH, int output);
int chaos_server_release(void* server, long long job_id);

/* Seeded mixed-job stream: Poisson arrivals at rate lambda for duration_s,
 * kinds uniform over {bifurcation 1-D Rossler, LLE Rossler, basins Lorenz},
 * sizes log-uniform in [*_min, *_max] points, physics of the benchmark
 * configurations with every horizon multiplied by horizon_scale.  mode 0
 * serves them through the engine, mode 1 runs each job alone with the
 * one-shot entry points in arrival order.  Latency = completion - arrival. */
typedef struct chaos_stream_config {
    double lambda, duration_s, horizon_scale;
    uint64_t seed;
    int bif_min, bif_max, lle_min, lle_max, basins_min, basins_max;
    int mode;
    int max_jobs;           /* cap on generated jobs (0 = none)               */
    double progress_s;      /* progress line on stderr every progress_s of
                               wall time (0 = none)                          */
} chaos_stream_config;

typedef struct chaos_stream_report {
    int jobs;
    int jobs_kind[3];
    double makespan_s, rk4_steps, steps_per_s, jobs_per_s;
    double p50_ms, p99_ms, mean_ms, max_ms;
    double p99_kind_ms[3];
    double busy_ms;         /* engine mode: device time of engine launches    */
    long long launches;
    double finalize_ms;     /* engine mode: host wall time finishing jobs     */
} chaos_stream_report;

int chaos_stream_config_default(chaos_stream_config* cfg);
/* The jobs of the stream cfg describes, in arrival order (what
 * chaos_stream_run serves): jobs[k] and arrival_ms[k] for k < cap (either may
 * be NULL); *count receives the number of jobs in the stream. */
int chaos_stream_jobs(const chaos_stream_config* cfg, chaos_job* jobs, double* arrival_ms, int cap,
                      int* count);
int chaos_stream_run(int device, const chaos_stream_config* cfg,
                     const chaos_engine_params* p, chaos_stream_report* rep,
                     double* latencies_ms /* [jobs] or NULL */, int latencies_cap);

/* Autotuner: a coordinate search over steps per launch, points per tile,
 * waves per SIMD, block size and graph replay.  Each candidate serves the
 * same short seeded stream (seed 7, 200 jobs/s for 1.5 s, the served kinds at
 * full horizons) and is scored by its power, RK4 steps/s over the mean job
 * latency; candidates within 1 % of the best power are ties, and the one of
 * them with the highest rate wins.  Writes it to *best (*best_steps_per_s:
 * its rate) and, if cache_path is not NULL, to that file with every
 * candidate's figures (read back by chaos_autotune_load). */
int chaos_autotune(int device, const char* cache_path, chaos_engine_params* best,
                   double* best_steps_per_s);
int chaos_autotune_load(const char* cache_path, chaos_engine_params* p);

#ifdef __cplusplus
}
#endif
#endif /* CHAOSLIBRARY_H */
