/*
 * fgx.h -- C ABI of the MI355X-native FARGO-style polar-grid disk engine.
 *
 * The hot path is the reference's per-DT hydro step AlgoGas
 * (src/SourceEuler.cu:208-301) with everything it calls.  The reference
 * exposes it as process-global state driven by main() (src/Main.cu:87-401);
 * this header replaces that surface with handle-based entry points.  Each
 * entry point below names the reference function(s) it stands in for.
 *
 * Conventions
 *   - Every function returns FGX_OK (0) or a negative FGX_E* code; the text of
 *     the last error on the calling thread is available from fgx_last_error().
 *     Nothing here calls exit(): the reference's gpuErrchk/exit paths
 *     (include/Kernels.cuh:5-13, src/Interpret.cu:173-197) become codes.
 *   - Plain C types only.  Field buffers are host memory; the library copies.
 *   - A field of a disk holds (nrad+1)*nsec values, ring-major (index i*nsec+j,
 *     i = ring, j = sector); ring nrad is the reference's ghost ring
 *     (src/Main.cu:199).
 *   - A handle is not re-entrant; different handles may be driven from
 *     different threads.
 */
#ifndef FGX_H
#define FGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGX_VERSION 100

/* ---- status codes ---------------------------------------------------- */
#define FGX_OK          0
#define FGX_EPARAM     -1   /* invalid parameter / parse error            */
#define FGX_EDEVICE    -2   /* HIP runtime error                          */
#define FGX_ENUMERIC   -3   /* non-finite or negative state detected       */
#define FGX_EIO        -4   /* file could not be read or written          */
#define FGX_ENOMEM     -5   /* allocation failed / over the memory budget */
#define FGX_ERANGE     -6   /* index or size out of range                 */

/* ---- arithmetic of the fields ------------------------------------------ */
#define FGX_F32 0           /* FP32 fields (the reference's precision)     */
#define FGX_F64 1           /* FP64 fields (verification mode)             */

#define FGX_MAX_PLANETS 8

/* ---- run parameters: the reference's .par keys and derived flags ---------
 * Keys and defaults follow include/Param.cuh:9-61; derived flags follow
 * src/Interpret.cu:126-202.  Floating keys are stored as float, as the
 * reference stores them.  Fields after `ext_` are keys the reference does not
 * know (it ignores unknown keys, src/Interpret.cu:35-95), so .par files that
 * use them still run under the reference. */
typedef struct fgx_params {
  int   nrad, nsec, ninterm, ntot;
  float dt, sigma0, masstaper, rmin, rmax, thicknesssmoothing, rochesmoothing,
        aspectratio, viscosity, alphaviscosity, sigmaslope, releaseradius,
        releasedate, omegaframe, imposeddiskdrift, flaringindex, eccentricity,
        cavityradius, cavityratio, cavitywidth, transitionradius,
        transitionratio, transitionwidth, lambdadoubling, adiabaticindex,
        coolingtime0;
  char  outputdir[512], planetconfig[512], innerboundary[64], labeladvection[64],
        transport[64], radialspacing[64], disk[64], frame[64],
        outersourcemass[64], writedensity[64], writevelocity[64],
        writeenergy[64], writetemperature[64], writedivv[64], writeqplus[64],
        indirectterm[64], excludehill[64], selfgravity[64], cicplanet[64],
        forcedcircular[64], zmplus[64], adiabatic_s[64], cooling_s[64];
  /* derived flags (YES=1 / NO=0), src/Interpret.cu:23-31,126-202 */
  int   log_grid, fast_transport, open_inner, nonreflecting, evanescent,
        is_disk, corotating, guiding_center, indirect_term, self_gravity,
        sg_zero_mode, zmplus_f, adiabatic, cooling, exclude_hill, cic_planet,
        forced_circular, viscosity_alpha, roche_smoothing, advect_label,
        outer_source_mass, write_density, write_velocity, write_energy,
        write_temperature, write_divv, write_qplus;
  /* command-line switches of src/Main.cu:108-181 that affect the step */
  int   sloppy_cfl;            /* -c */
  int   centrifugal_balance;   /* -b */
  float scaling_factor;        /* -f */
  /* ext_: keys unknown to the reference ------------------------------- */
  double ext_noise_amplitude;  /* NOISEAMPLITUDE: Sigma *= 1 + A(2u-1)       */
  uint64_t ext_noise_seed;     /* NOISESEED (default 20261203)              */
  int   ext_disk_id;           /* DISKID: added to the seed                 */
  int   ext_substeps;          /* SUBSTEPS: >0 forces that many equal        */
                               /* substeps per DT instead of ConditionCFL    */
  int   ext_fix_taurp_corner;  /* QUIRKFIXTAURP: corner density uses j-1     */
  int   ext_fix_pp_distance;   /* QUIRKFIXPPDIST: planet-planet |ri-rj|^3    */
  int   ext_fix_sg_wrap;       /* QUIRKFIXSGWRAP: periodic j-1 in SG update  */
  float ext_sg_softening;      /* SGSOFTENING: >0 overrides SG epsilon/H     */
  int   ext_gas_names;         /* GASNAMES: 0 dens/ layout, 1 gasdens*.dat, 2 both */
  int   ext_checkpoint;        /* CHECKPOINT: fgx-run writes restart<N>.bin at each output */
  int   ext_max_substeps;      /* MAXSUBSTEPS: a disk whose CFL count in a DT exceeds it is   */
                               /* stopped as numerically unhealthy (default 100000)       */
  int   ext_stubbed_physics;   /* STUBBEDPHYSICS: run what the reference parses but leaves  */
                               /* stubbed, with FARGO's semantics: planet accretion     */
                               /* (src/Planet.cu:7-47), label advection                 */
                               /* (src/TransportEuler.cu:28,34,68), -b centrifugal      */
                               /* balance (src/Pframeforce.cu:202-237) and the outer    */
                               /* mass source (src/SideEuler.cu:143-146,181)            */
} fgx_params;

/* ---- planetary system: the reference's PlanetarySystem (include/Types.cuh:2-15) */
typedef struct fgx_planets {
  int    nb;
  double mass[FGX_MAX_PLANETS], x[FGX_MAX_PLANETS], y[FGX_MAX_PLANETS],
         vx[FGX_MAX_PLANETS], vy[FGX_MAX_PLANETS], acc[FGX_MAX_PLANETS];
  int    feel_disk[FGX_MAX_PLANETS], feel_others[FGX_MAX_PLANETS];
  char   name[FGX_MAX_PLANETS][64];
} fgx_planets;

/* ---- fields and ring profiles readable from a disk ---------------------- */
enum fgx_field {
  FGX_DENS = 0, FGX_VRAD = 1, FGX_VTHETA = 2, FGX_ENERGY = 3,
  FGX_TEMPERATURE = 4, FGX_SOUNDSPEED = 5, FGX_DIVV = 6, FGX_QPLUS = 7,
  FGX_LABEL = 8, FGX_POTENTIAL = 9,
  FGX_SG_ACCR = 10, FGX_SG_ACCT = 11,   /* self-gravity accelerations SG_Accr, SG_Acct (SgMain.cu) */
  FGX_NFIELDS = 12
};
enum fgx_ring {
  FGX_R_RADII = 0, FGX_R_RMED = 1, FGX_R_SURF = 2, FGX_R_SIGMAMED = 3,
  FGX_R_ENERGYMED = 4, FGX_R_VISCOSITY = 5, FGX_R_VMED = 6, FGX_R_NSHIFT = 7,
  FGX_R_CS = 8, FGX_R_NEWDT = 9, FGX_R_INVRMED = 10, FGX_NRINGS = 11
};

/* per-disk scalar state (host copy) */
typedef struct fgx_disk_info {
  int    nrad, nsec, precision;
  double physical_time;        /* PhysicalTime (kept in FP64, see quirk Q14) */
  double omega_frame;          /* OmegaFrame                                  */
  double lost_mass;            /* LostMass (src/TransportEuler.cu:17,71)      */
  double mdcp1, exces_mdcp;    /* src/SourceEuler.cu:294-295                  */
  double last_dt;              /* dt of the last substep                      */
  int    last_ncfl;            /* last gastimestepcfl                         */
  long long substeps;          /* substeps completed since creation           */
  long long dt_steps;          /* AlgoGas calls completed since creation      */
  int    health;               /* 0 ok, else FGX_HEALTH_* flags: the disk was   */
                               /* stopped and frozen at the substep that set them */
} fgx_disk_info;

/* numeric health of a disk (no counterpart in the reference, whose crash
 * detection is disabled, src/SourceEuler.cu:62-74, 268-276).  A disk that
 * sets any flag is frozen where its substep starts; the other disks of its
 * batch go on, and fgx_batch_step / fgx_disk_step return FGX_ENUMERIC. */
#define FGX_HEALTH_DENSITY   1  /* a density not positive or not a number      */
#define FGX_HEALTH_NAN       2  /* a velocity or sound speed not a number      */
#define FGX_HEALTH_TIMESTEP  4  /* the CFL time step not positive             */
#define FGX_HEALTH_SUBSTEPS  8  /* the CFL count of a DT above MAXSUBSTEPS     */

typedef struct fgx_batch fgx_batch;
typedef struct fgx_batch fgx_disk;   /* a disk is a batch of one */

/* ---- library --------------------------------------------------------- */
int         fgx_version(void);
const char* fgx_last_error(void);
/* sizeof(fgx_params), sizeof(fgx_planets): lets bindings check layouts */
size_t      fgx_sizeof_params(void);
size_t      fgx_sizeof_planets(void);
int         fgx_device_count(int* n);

/* ---- parameters (src/Interpret.cu, include/Param.cuh) ------------------- */
/* defaults of include/Param.cuh:9-61 and src/Interpret.cu:23-31 */
int fgx_params_default(fgx_params* p);
/* ReadVariables (src/Interpret.cu:97-208) from a file / from text; the
 * derived-flag pass and its validity checks are included */
int fgx_params_load(const char* par_path, fgx_params* p);
int fgx_params_parse(const char* text, fgx_params* p);
/* SearchVariable (src/Interpret.cu:35-95) for one key followed by the
 * derived-flag pass; key is case-insensitive, value is raw */
int fgx_params_set(fgx_params* p, const char* key, const char* value);
/* value of one key as text ("%.17g" for numbers) */
int fgx_params_get(const fgx_params* p, const char* key, char* out, size_t n);

/* ---- planets (src/Psys.cu:85-134) ----------------------------------------
 * InitPlanetarySystem: CICPlanet placement needs the grid, so the params are
 * passed; ECCENTRICITY comes from them as well. */
int fgx_planets_load(const char* cfg_path, const fgx_params* p, fgx_planets* out);
int fgx_planets_parse(const char* text, const fgx_params* p, fgx_planets* out);

/* ---- one disk -----------------------------------------------------------
 * fgx_disk_create = everything main() does before its loop
 * (src/Main.cu:182-344): grid (FillPolar1DArrays, src/SourceEuler.cu:76-173),
 * InitGasDensity/InitGasEnergy (src/Pframeforce.cu:173-192), self-gravity
 * start-up (src/Main.cu:302-329), InitEuler (src/SourceEuler.cu:177-204).
 * Optional seeded density noise (ext_noise_amplitude) is applied after
 * InitGasDensity. precision is FGX_F32 or FGX_F64. */
int fgx_disk_create(const fgx_params* p, const fgx_planets* pl, int precision,
                    fgx_disk** out);
/* n_dt calls of AlgoGas (src/SourceEuler.cu:208-301) */
int fgx_disk_step(fgx_disk* d, int n_dt);
int fgx_disk_destroy(fgx_disk* d);
int fgx_disk_info_get(fgx_disk* d, fgx_disk_info* info);
/* count >= nrad*nsec; with count >= (nrad+1)*nsec the ghost ring is included.
 * dtype FGX_F32 / FGX_F64 is the host buffer's type. */
int fgx_disk_read_field(fgx_disk* d, int field, int dtype, void* host, size_t count);
int fgx_disk_write_field(fgx_disk* d, int field, int dtype, const void* host, size_t count);
int fgx_disk_read_ring(fgx_disk* d, int ring, int dtype, void* host, size_t count);
int fgx_disk_planets(fgx_disk* d, fgx_planets* out);
int fgx_disk_set_planets(fgx_disk* d, const fgx_planets* pl);

/* ---- output files (src/Output.cu, src/Force.cu:76-138) -------------------- */
/* SendOutput (src/Output.cu:155-170) for output number `index` */
int fgx_disk_write_outputs(fgx_disk* d, int index);
/* WriteBigPlanetSystemFile (src/Output.cu:90-102) + UpdateLog
 * (src/Force.cu:76-138) for time step `timestep` */
int fgx_disk_write_logs(fgx_disk* d, int timestep);
/* WriteDim (src/Output.cu:137-153), used_rad.dat (src/SourceEuler.cu:159-171),
 * EmptyPlanetSystemFile (src/Output.cu:12-32) */
int fgx_disk_write_setup_files(fgx_disk* d);
/* disk->point force, all eleven Hill-cut variants (ComputeForce,
 * src/Force.cu:31-67): out[44] laid out as the reference's GlobalForce
 * (fxi[0..10], fxo[0..10], fyi[0..10], fyo[0..10]) */
int fgx_disk_force(fgx_disk* d, double x, double y, double smoothing, double mass,
                   double out[44]);
/* fgx_disk_force with no allocation of any kind: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The kernels'
 * per-ring partials and the totals live in a static device array of the
 * library (4096 rings x 44 doubles + 64, about 1.4 MB per GPU, held until the
 * process ends); a disk of more than 4096 rings is refused with FGX_ERANGE.
 * The result equals fgx_disk_force's bit for bit.  Calls on one GPU share
 * that array: calls that overlap in time on one GPU race. */
int fgx_disk_force_noalloc(fgx_disk* d, double x, double y, double smoothing, double mass,
                           double out[44]);

/* ---- batches of many disks (no reference counterpart: the reference runs
 * one disk per process, src/Main.cu:87-98).  All disks of one batch share
 * the mode flags (adiabatic, boundary kind, transport, self-gravity) and may
 * differ in grid size and every numeric parameter. */
int fgx_batch_create(int n, const fgx_params* p, const fgx_planets* pl,
                     int precision, fgx_batch** out);
int fgx_batch_step(fgx_batch* b, int n_dt);
int fgx_batch_size(fgx_batch* b, int* n);
int fgx_batch_destroy(fgx_batch* b);
int fgx_batch_info(fgx_batch* b, int disk, fgx_disk_info* info);
/* planets of disk `disk` (positions, velocities, masses: accretion changes
 * them with StubbedPhysics), as fgx_disk_planets for one disk */
int fgx_batch_planets(fgx_batch* b, int disk, fgx_planets* out);
int fgx_batch_read_field(fgx_batch* b, int disk, int field, int dtype, void* host,
                         size_t count);
int fgx_batch_write_field(fgx_batch* b, int disk, int field, int dtype,
                          const void* host, size_t count);
/* freeze (1) or thaw (0) disk `disk` of a batch: later fgx_batch_step calls
 * leave a frozen disk untouched (its tiles are skipped).  No counterpart in
 * the reference; the serving loop retires finished requests with it. */
int fgx_batch_freeze(fgx_batch* b, int disk, int frozen);
/* Complete evolving state of disk `disk` (fields in the engine's precision,
 * vtheta as its residual, planets, clocks, frame, ring profiles carried between
 * substeps) to `path`, between DTs; fgx_batch_restore loads it into a disk of
 * the same parameters, which then continues bit-identically.  The reference
 * parses a restart switch but never restores anything (-s n, src/Main.cu:147-155,
 * 175-180); fgx-run -s n restores from OUTPUTDIR/restart<n>.bin, written at each
 * output when CHECKPOINT is YES. */
int fgx_batch_checkpoint(fgx_batch* b, int disk, const char* path);
int fgx_batch_restore(fgx_batch* b, int disk, const char* path);
/* Slot `disk` of a live batch takes a new disk: parameters and planets of
 * the same NRAD x NSEC grid and batch mode flags, set up exactly as
 * fgx_batch_create would set it up (it then evolves bit-identically to that
 * disk alone).  Nothing is allocated; the other disks keep their state and
 * the frozen flag of the slot is cleared.  FGX_EPARAM when the grid or a mode
 * differs, or when the disk needs a pass (accretion, outer mass source,
 * viscosity) the batch was created without.  No counterpart in the reference,
 * which runs one disk per process (src/Main.cu:87-401); the serving loop
 * refills the slots of finished requests with it. */
int fgx_batch_reset_disk(fgx_batch* b, int disk, const fgx_params* p, const fgx_planets* pl);
/* bytes of device memory a disk of these parameters needs */
int fgx_footprint_bytes(const fgx_params* p, int precision, size_t* bytes);

/* ---- tile shapes and their autotuner (no counterpart in the reference, whose
 * kernels launch one thread per cell, src/SourceEuler.cu:208-301)
 * The batch kernels walk strips of rings; how tall the strips are, and below
 * how many active cells a substep switches to the short-strip tiling, set how
 * well a batch fills the device.  The fields do not depend on the tiling: any
 * tiling gives the same bits.  A zero field means the engine's own choice for
 * the batch (and then an FGX_STRIP_RINGS / FGX_SOURCE_RINGS /
 * FGX_FINE_SOURCE_RINGS / FGX_FINE_CELLS environment variable, if set). */
typedef struct fgx_tiling {
  int    strip_rings;        /* radial sweep strips: rings per strip, 2..64       */
  int    source_rings;       /* fused source pass strips: rings, 1..24            */
  int    fine_source_rings;  /* source strips of the short-strip tiling, 1..source_rings */
  int    fine_strip_rings;   /* radial strips of the short-strip tiling, 1..strip_rings */
  double fine_cells;         /* short strips below this many active cells; <0: never */
} fgx_tiling;
typedef struct fgx_tune_report {
  int    from_cache;     /* 1: the tiling came from the cache file, nothing was timed */
  int    candidates;     /* tilings timed                                        */
  double default_rate;   /* cell-updates/s of the engine's own tiling            */
  double best_rate;      /* cell-updates/s of the tiling returned                 */
  double seconds;        /* wall time the tuning took                            */
  char   signature[128]; /* the workload key the choice is recorded under        */
} fgx_tune_report;
/* fgx_batch_create with a tiling (NULL: the engine's choice) */
int fgx_batch_create_tiled(int n, const fgx_params* p, const fgx_planets* pl, int precision,
                           const fgx_tiling* t, fgx_batch** out);
/* the tiling a batch runs with, every field resolved */
int fgx_batch_tiling(fgx_batch* b, fgx_tiling* out);
/* Autotuner: the tiling under which this batch of disks advances fastest.
 * Scratch batches of the same disks are created with candidate tilings (a
 * coordinate search from the engine's own choice over the four strip heights
 * and the short-strip threshold) and timed over n_dt DTs after one warm-up DT;
 * a candidate replaces the incumbent only when it is more than 1% faster.
 * The choice is recorded in cache_path (one line per workload signature:
 * precision, disk shapes, mode flags and the device name); a later call with a
 * matching signature reads it back without timing anything.  cache_path may be
 * NULL (no cache).  Needs device memory for one scratch batch at a time. */
int fgx_autotune(int n, const fgx_params* p, const fgx_planets* pl, int precision, int n_dt,
                 const char* cache_path, fgx_tiling* out, fgx_tune_report* report);

/* ---- disk groups: one set of disks as several concurrent batches (no
 * counterpart in the reference, which runs one disk per process,
 * src/Main.cu:87-98).  The disks are dealt, largest first, to `parts`
 * batches (each to the one with the fewest cells so far; a batch keeps its
 * disks in their given order), all of the tiling t (NULL: the engine's
 * choice per batch).  fgx_group_step advances every batch n_dt DTs, each on
 * its own stream and host thread, so one batch's sweeps fill the CUs the
 * other's short per-disk launches leave idle; it returns when all are done.
 * A disk's fields are those it has in one batch of all the disks, bit for
 * bit.  Its batch and its index there (fgx_group_disk) serve every
 * fgx_batch_* call on it (read, write, info, planets, freeze, checkpoint);
 * those batches belong to the group, which destroys them. */
#define FGX_GROUP_MAX_PARTS 8
typedef struct fgx_group fgx_group;
int fgx_group_create(int n, const fgx_params* p, const fgx_planets* pl, int precision, int parts,
                     const fgx_tiling* t, fgx_group** out);
int fgx_group_step(fgx_group* g, int n_dt);
int fgx_group_parts(fgx_group* g, int* n);
int fgx_group_batch(fgx_group* g, int part, fgx_batch** out);
int fgx_group_disk(fgx_group* g, int disk, fgx_batch** batch, int* index);
int fgx_group_destroy(fgx_group* g);

/* ---- serving loop (a stream of run requests; no counterpart in the reference)
 * Requests (one disk each: parameters, planets, number of DTs, arrival time
 * in seconds from the start of fgx_server_run) are served by continuous
 * batching.  At most max_live_batches batches are live, each advanced one DT
 * at a time on its own stream and host thread.  A new batch holds a group of
 * waiting requests that share the batch mode flags, plus spare_slots frozen
 * slots of the group's grids, while the summed device footprint of the live
 * batches stays within hbm_budget.  A request that has done its DTs is frozen
 * and its slot returns to the pool at once: at the batch's next DT boundary
 * a waiting request of the same grid and modes takes it (fgx_batch_reset_disk,
 * no allocation).  A batch with every slot free is released when a waiting
 * request needs its memory or place, or when nothing is left to come.  Every
 * snapshot_every DTs of a request its fields are copied asynchronously into
 * pinned host memory (the reference writes its outputs every NINTERM steps,
 * src/Main.cu:348-371).  A request's fields evolve exactly as they would
 * alone. */
typedef struct fgx_server fgx_server;
typedef struct fgx_server_config {
  int    precision;        /* FGX_F32 / FGX_F64                                */
  size_t hbm_budget;       /* bytes of device memory the live batches may use  */
  int    max_batch_disks;  /* most requests in one batch                       */
  int    snapshot_every;   /* DTs between snapshots (0: none)                  */
  double batch_window;     /* seconds to wait for more arrivals to batch with  */
  int    keep_final;       /* keep each request's final fields for read-back   */
  int    max_live_batches; /* batches live at once (0: 4, the hardware queues)  */
  int    spare_slots;      /* frozen slots a new batch holds beside its group   */
} fgx_server_config;
typedef struct fgx_request_info {
  double arrival, admit, finish;      /* seconds from the start of the run   */
  long long substeps, cell_updates;   /* CFL substeps done, cells x substeps */
  int    n_dt, snapshots, batch, batch_size, status;   /* status: FGX_OK or error */
  int    slot;                        /* the slot of its batch it ran in     */
} fgx_request_info;
typedef struct fgx_server_stats {
  int    batches;       /* batches created                                 */
  int    refills;       /* requests set up in a freed slot of a live batch */
  int    peak_live;     /* most batches live at once                       */
  size_t peak_bytes;    /* most device memory the live batches held        */
} fgx_server_stats;
int fgx_server_create(const fgx_server_config* cfg, fgx_server** out);
int fgx_server_submit(fgx_server* s, const fgx_params* p, const fgx_planets* pl, int n_dt,
                      double arrival, int* id);
int fgx_server_run(fgx_server* s);      /* serves every submitted request    */
int fgx_server_request(fgx_server* s, int id, fgx_request_info* out);
int fgx_server_read_field(fgx_server* s, int id, int field, int dtype, void* host, size_t count);
int fgx_server_stats_get(fgx_server* s, fgx_server_stats* out);
int fgx_server_destroy(fgx_server* s);

/* ---- profiling: average device time per kernel class ---------------------
 * enable=1 brackets each launch with HIP events on the library's stream. */
int fgx_profile_enable(fgx_batch* b, int enable);
/* name of kernel class k (0..) or NULL past the end, and its totals */
int fgx_profile_get(fgx_batch* b, int k, const char** name, double* total_ms,
                    long long* launches, double* bytes_per_launch);
int fgx_profile_reset(fgx_batch* b);

/* ---- the reference executable: fargoGPU [flags] file.par (src/Main.cu:87-401) */
int fgx_run_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif /* FGX_H */
