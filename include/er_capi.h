/*
 * er_capi.h -- C ABI of the MI355X erdst hot path (per-frame solute-solvent
 * energies + energy-coordinate distributions).
 *
 * The reference exposes this path only as Fortran module procedures that
 * engconst/engstore call on shared module arrays.  Each entry point below
 * names the reference procedure(s) it replaces, so a Fortran host binds it
 * with iso_c_binding (fortran/er_capi_mod.f90 in the package is that
 * interface module).
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero ER_E_* status on
 *    failure; er_last_error() then gives the reference's 7-character halt
 *    code (erdst/mpiproc.F90:184-230) plus the species and energy involved.
 *    The library never stops the process.
 *  - Host arrays are owned by the caller, copied in, and never retained
 *    past the call (the *_device variants are the documented exception).
 *  - Index arrays are 1-based and arrays are column-major, exactly as the
 *    Fortran module engmain holds them (erdst/engmain.F90:314-348).
 *  - A context is used by one host thread.  It owns one HIP stream and all
 *    device memory.  er_run() is asynchronous; er_sync(), er_fetch() and
 *    er_get_uvengy() wait for it.
 *  - Arithmetic: pair and PME energies in FP32 (the reference's default
 *    build); the reciprocal-grid self energy, the weights and every
 *    distribution accumulator (edens, ecorr, eself, engnorm, ...) in FP64.
 */
#ifndef ER_CAPI_H
#define ER_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct er_ctx er_ctx;

/* status codes */
enum {
    ER_OK = 0,
    ER_E_ARG = 1,        /* invalid argument / inconsistent sizes            */
    ER_E_UNSUPPORTED = 2,/* option outside what this build implements        */
    ER_E_HIP = 3,        /* HIP runtime error                                */
    ER_E_HALT = 4,       /* a reference halt condition (see er_last_error)   */
    ER_E_STATE = 5       /* call out of order (e.g. er_run before a frame)   */
};

/* run modes for er_run */
/* ER_PLAIN: one insertion per launch sequence, the host waiting for each
 * (the reference's per-insertion host loop); ER_PLAIN_ASYNC: the same
 * launches without the waits; ER_BATCHED: every insertion of the frame in one
 * launch per kernel class.  All three give the same results to rounding. */
enum { ER_PLAIN = 0, ER_BATCHED = 1, ER_PLAIN_ASYNC = 2 };

/* Topology of the physical system, as setparam leaves it in module engmain
 * (erdst/setconf.F90:516-844).  "Physical" = molecules whose coordinates are
 * in the trajectory (solvent, plus the solute molecules in soln mode). */
typedef struct {
    int32_t numatm;             /* atoms of the physical system              */
    int32_t nummol;             /* molecules of the physical system (slvmax) */
    int32_t ljtype_max;
    int32_t numslv;             /* number of solvent species                 */
    const int32_t *mol_begin_index; /* [nummol+1], 1-based                   */
    const int32_t *ljtype;      /* [numatm], 1-based                         */
    const int32_t *uvspec;      /* [nummol]: species 1..numslv, 0 = solute   */
    const float *charge;        /* [numatm], already x 18.22261721           */
    const float *ljene_mat;     /* [ljtype_max^2], column-major              */
    const float *ljlensq_mat;   /* [ljtype_max^2], column-major              */
} er_topology;

/* The &ene_param values the hot path reads (erdst/engmain.F90:376-388),
 * after iniparam's defaulting (erdst/setconf.F90:133-438). */
typedef struct {
    float elecut, lwljcut, upljcut, screen;
    int32_t ljswitch;           /* 0..3, erdst/engmain.F90:358              */
    int32_t cltype;             /* 0 bare Coulomb, 2 PME, 3 PPPM (even splodr) */
    int32_t splodr;             /* B-spline order, 3..10                    */
    int32_t ms[3];              /* ms1max, ms2max, ms3max: any size from
                                   max(16, 2 splodr) to 512, the plane
                                   ms2max x ms3max up to about 128 x 128    */
    int32_t boxshp;             /* 1 periodic; 0 open (cltype 0, estype 1)  */
    int32_t slttype;            /* 1 soln, 2 refs rigid, 3 refs flexible    */
    int32_t wgtslf, estype, corrcal, selfcal;
    double temp;                /* kT in kcal/mol (setconf.F90:304)         */
    int32_t force_calculation;  /* 1: the 'mbin' warning does not stop the run
                                   (engmain force_calculation, mpiproc.F90:238-252) */
} er_params;

/* Replaces: setparam's device uploads (setconf.F90:656-841), recpcal_init
 * (recpcal.F90:40-95), enginit's ecorr allocation (engproc.F90:283-287). */
int er_create(const er_topology *topo, const er_params *prm, er_ctx **out);

/* Replaces: enginit's mesh (engproc.F90:127-281).  uvmax/uvsoft: [numslv];
 * uvcrd: [sum(uvmax)] FP64 concatenated per species; escrd: [esmax] (only
 * read when selfcal = 1). */
int er_set_mesh(er_ctx *c, const int32_t *uvmax, const int32_t *uvsoft,
                const double *uvcrd, int32_t esmax, const double *escrd);

/* Test particle of a refs run: the last molecule in the reference's system
 * (engproc.F90:99-101).  ljtype 1-based, charge x 18.22261721. */
int er_set_solute(er_ctx *c, int32_t nsite, const int32_t *ljtype, const float *charge);

/* Several test particles analysed over the same frames in one context (refs
 * runs).  The reference runs erdst once per solute (one test particle per
 * run, engproc.F90:99-101); here each solute k = 1..nsolute is a slot with
 * its own insertions, energies and block accumulators, and one er_run
 * processes every slot that has insertions, sharing the frame's
 * preparation.  nsite: [nsolute]; ljtype (1-based) and charge (x 18.22261721):
 * concatenated in solute order.  Resets every slot's accumulators; slot 1
 * is selected afterwards. */
int er_set_solutes(er_ctx *c, int32_t nsolute, const int32_t *nsite, const int32_t *ljtype,
                   const float *charge);

/* Selects slot k (1-based) for the calls that address one solute:
 * er_set_solute, er_set_insertions[_device], er_get_uvengy, er_fetch.
 * er_clear clears every slot; er_sync reports the first slot that halted. */
int er_select_solute(er_ctx *c, int32_t k);

/* Solute molecules of a soln run (sltlist, engproc.F90:72-96), 1-based. */
int er_set_solute_molecules(er_ctx *c, int32_t nslt, const int32_t *tagslt);

/* Replaces: the frame arrays of getconf_parallel (setconf.F90:849-986) that
 * engconst consumes.  sitepos: [3*numatm] column-major; cell: [9]
 * column-major (column j = box vector j). */
int er_set_frame(er_ctx *c, const float *sitepos, const float *cell,
                 float stat_weight_system);
/* Same, but sitepos is a device pointer that stays valid until the next
 * er_run() has completed (no copy is made). */
int er_set_frame_device(er_ctx *c, const float *sitepos_dev, const float *cell,
                        float stat_weight_system);

/* Replaces: instslt('proc') x maxdst (engproc.F90:810-812) -- the caller
 * places the insertions.  sltpos: [3*nsite*nins] (insertion-major, i.e. the
 * sitepos extension slots of erdst/setconf.F90:607-611); stat_weight_solute:
 * [nins] or NULL for all 1. */
int er_set_insertions(er_ctx *c, int32_t nins, const float *sltpos,
                      const float *stat_weight_solute);
int er_set_insertions_device(er_ctx *c, int32_t nins, const float *sltpos_dev,
                             const float *stat_weight_solute);

/* Replaces: the body of engconst for one frame (engproc.F90:419-458):
 * realcal_prepare + recpcal_spline_greenfunc + recpcal_prepare_solvent +
 * get_uv_energy_refs|soln + update_histogram x maxdst.  Results accumulate
 * into the block accumulators.  ER_PLAIN / ER_PLAIN_ASYNC launch one
 * insertion at a time; ER_BATCHED runs every insertion of the frame, of every solute that has
 * insertions (er_set_solutes), in one launch per kernel class (the
 * reciprocal part in chunks of insertions, one chunk by default when the
 * tuner picks it). */
int er_run(er_ctx *c, int32_t mode);

/* No counterpart in the reference: er_run for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating library routine: every per-run buffer of the context must
 * already hold this run's shapes (an earlier er_run of at least as many
 * insertions per solute, same launch shapes, sizes them; otherwise
 * ER_E_STATE), and the split-K partial sums of edens/ecorr and the weight
 * partials live in a 64 MiB array of the library's own that each GPU holds
 * until the process ends.  A run whose partial sums exceed it (about
 * nslices x ermax^2 x 8 bytes with ecorr on) returns ER_E_UNSUPPORTED.
 * Results equal er_run's.  All contexts on one GPU share that array, so
 * er_run_noalloc calls that overlap on one GPU race: issue them from one
 * stream at a time.  Profiling (er_profile) is suspended for the call. */
int er_run_noalloc(er_ctx *c, int32_t mode);

/* HIP graphs of whole runs; no counterpart in the reference, whose frame
 * loop (erdst/enganal.F90:86-95 -> engconst per frame) launches every
 * kernel from the host.  Between er_graph_begin and er_graph_end the calls
 * er_set_frame_device, er_select_solute, er_set_insertions_device (with
 * stat_weight_solute NULL) and er_run(ER_BATCHED) are recorded, not
 * executed; any other call on the context fails with ER_E_STATE and spoils
 * the recording.  er_graph_end instantiates the recording and returns a
 * handle >= 1; er_graph_launch replays it with one launch on the context's
 * stream, asynchronously (er_sync waits), adding to the block accumulators
 * exactly what the recorded calls would add.
 *  - The graph keeps the device pointers, cells, weights and launch shapes
 *    of the recording: the caller refills the same device buffers between
 *    replays.  A replay needs the same solutes, insertion counts and launch
 *    shapes as the recording; a run of them before er_graph_begin sizes the
 *    buffers (nothing is allocated while recording).
 *  - er_graph_launch refuses (ER_E_STATE) a graph whose device buffers were
 *    reallocated since, or whose solutes, shapes or influence-function cell
 *    differ from the context's; record it again then.
 *  - Halts and warnings of replayed runs are reported by er_sync as for
 *    er_run.  A typical use records one batch of frames (16 frames of 1000
 *    insertions) and replays it per batch. */
int er_graph_begin(er_ctx *c);
int er_graph_end(er_ctx *c, int32_t *graph);
int er_graph_launch(er_ctx *c, int32_t graph);
int er_graph_destroy(er_ctx *c, int32_t graph);

/* Waits for the context's stream; returns ER_E_HALT if a run hit a halt
 * condition (getiduv's 'eng_min' / 'eng_sft', engproc.F90:1142-1156) or,
 * unless force_calculation is set, the 'mbin' warning: an energy below
 * 1e3 kcal/mol in a species' open top bin (engproc.F90:1158-1167), on which
 * the reference stops (warning, mpiproc.F90:238-252).  The earliest event in
 * (run, insertion, molecule) order is reported, as the reference meets
 * them; er_last_error gives its code ("mbin" for the warning), species and
 * energy. */
int er_sync(er_ctx *c);

/* Per-insertion energies of the last er_run, for parity checks.
 * uvengy: [nins*nummol] (= uvengy(slvmax, maxdst) column-major); uvengy0:
 * [nins].  Either pointer may be NULL.  nins is the selected solute's
 * insertion count of that run (er_get_nins); the arrays must hold it. */
int er_get_uvengy(er_ctx *c, float *uvengy, float *uvengy0);

/* No counterpart in the reference, whose uvengy array has the fixed bound
 * maxdst: the row count of er_get_uvengy, i.e. the selected solute's
 * insertions in the last er_run (0 before any run). */
int er_get_nins(er_ctx *c, int32_t *nins);

/* Replaces: the data engstore reads (engproc.F90:520-725) before its
 * normalisation.  edens [ermax], ecorr [ermax*ermax] (may be NULL),
 * eself [esmax] (may be NULL), scal = {engnorm, engsmpl, avslf, voffset},
 * minuv/maxuv [numslv+1] (index 0 = solute self energy), slnuv [numslv]
 * (soln; may be NULL). */
int er_fetch(er_ctx *c, double *edens, double *ecorr, double *eself, double *scal,
             float *minuv, float *maxuv, double *slnuv);

/* Per-solute species sums of the last soln er_run (flceng, engproc.F90:918-932);
 * flceng: [numslv*nslt]. */
int er_get_flceng(er_ctx *c, double *flceng);

/* Replaces: engclear (engproc.F90:321-340). */
int er_clear(er_ctx *c);

/* code: 8 bytes, NUL-terminated 7-character halt code ("" if none). */
int er_last_error(er_ctx *c, char *code, int32_t *species, float *energy);

/* Human-readable text of the last failure (static storage of the context). */
const char *er_error_string(er_ctx *c);

void er_destroy(er_ctx *c);

/* Diagnostics, no counterpart in the reference: per-kernel device time of
 * the runs issued while enabled, measured with HIP events recorded on the
 * context's stream around each launch.  er_profile_read waits for the stream
 * and returns (then resets) the accumulated time and launch count of one
 * kernel class. */
enum { ER_K_PREP = 0, ER_K_SPEC2D, ER_K_AXIS1, ER_K_GATHER, ER_K_PAIRS, ER_K_HIST, ER_K_WEIGHTS,
       ER_K_ECORR, ER_K_COUNT };
int er_profile(er_ctx *c, int32_t enable);
int er_profile_read(er_ctx *c, int32_t kernel, double *total_ms, int64_t *launches);

/* Launch-shape tuning, no counterpart in the reference (its launch shapes
 * are fixed by the OpenACC compiler).  er_autotune times candidate shapes on
 * the frame and insertions already set, with HIP events on the context's
 * stream, over reps runs per candidate (0: 3):
 *   chunk       insertions per launch sequence: the current one, 256 ..
 *               4096 (whole batched runs of every solute);
 *   hist_slices histogram workgroups per 16 insertions of a refs run: the
 *               current one, 1 .. 64 (the histogram kernel's own time);
 *   pair_split  pair-search workgroups per insertion, per solute: the
 *               current one, 1 .. 64 (the pair kernel's own time);
 *   cell width  the pair search's least cell width (er_set_cell_width).
 * A candidate replaces the current shape only when it is more than 2%
 * faster, so the result is never worse than the current shapes by more than
 * timing noise; 0 stands for the built-in choice.  It restores every
 * solute's block accumulators, minuv/maxuv and halt state as they were, and
 * leaves the per-insertion energies of the current frame for
 * er_get_uvengy.  No shape changes any result, bit for bit (soln runs keep
 * the built-in histogram slices, which their flceng sums follow).
 * The shapes apply to ER_BATCHED runs; ER_PLAIN keeps its built-in ones.
 * er_set_tuning applies shapes kept from an earlier tuning (chunk 0 keeps
 * the current one; pair_split 0 and hist_slices 0 restore the built-in
 * choice; pair_split applies to the selected solute); er_get_tuning reads
 * them (pair_split of the selected solute). */
int er_autotune(er_ctx *c, int32_t reps);
int er_set_tuning(er_ctx *c, int32_t chunk, int32_t pair_split, int32_t hist_slices);
int er_get_tuning(er_ctx *c, int32_t *chunk, int32_t *pair_split, int32_t *hist_slices);
/* The least cell width (Angstrom) of the pair search's cell list, also tuned
 * by er_autotune (candidates 3.5, 5, 7 and 10; 0 = the built-in 3.5).  The
 * cells partition the molecules whatever their width, so no result changes;
 * the next run rebuilds the frame's cell list.  No counterpart in the
 * reference, which tests every (insertion, molecule) pair. */
int er_set_cell_width(er_ctx *c, float angstrom);
int er_get_cell_width(er_ctx *c, float *angstrom);

#ifdef __cplusplus
}
#endif
#endif
