/*
 * ddfgpu.h -- C ABI of the MI355X (gfx950) facet gridder / degridder.
 *
 * This library replaces the native module DDFacet.cbuild.Gridder._pyGridderSmearPols3x
 * (DDFacet/Gridder/GridderSmearPols.cc:349-370).  The Python shim
 * ddfacet_amd/_pyGridderSmearPols.py exports the reference names with the reference argument
 * lists and marshals them into the structs below; any other host (C, C++, Go over cgo, ...)
 * can call these entry points directly.
 *
 * Conventions
 *   - Every pointer inside an args struct is a DEVICE pointer (hipMalloc'd memory or a torch
 *     tensor's data_ptr()) unless the field comment says "host".
 *   - Entry points are asynchronous on `stream` (a hipStream_t passed as void*, NULL = the null
 *     stream) and return DDF_OK or a negative error code; ddf_last_error() returns the message of
 *     the last failure on the calling thread.  Argument errors are detected before any launch.
 *   - complex64 data are interleaved (re, im) float pairs; complex128 are (re, im) double pairs.
 *   - Correlation / Stokes codes are the casacore Stokes enum used by DDFacet
 *     (DDFacet/Data/ClassStokes.py:39-41): I=1 Q=2 U=3 V=4 RR=5 RL=6 LR=7 LL=8 XX=9 XY=10 YX=11 YY=12.
 */
#ifndef DDFGPU_H
#define DDFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDF_ABI_VERSION 1

enum ddf_status {
  DDF_OK = 0,
  DDF_EINVAL = -1,    /* invalid argument (reference: std::invalid_argument -> ValueError)   */
  DDF_ERUNTIME = -2,  /* runtime failure  (reference: std::runtime_error   -> RuntimeError)  */
  DDF_EHIP = -3,      /* HIP / hipFFT API failure                                            */
  DDF_ENOMEM = -4     /* device allocation failure                                           */
};

/* Per-facet w-projection convolution functions, the content of ClassWTermModified.Wplanes and
 * .WplanesConj (DDFacet/Imager/ModCF.py:560-593).  Plane p (0 <= p < nw) is Wplanes[p], plane
 * nw + p is WplanesConj[p]; both are complex64 [(S*overs)^2] in the reorganised layout
 * [overs][overs][S][S] (ModCF.py:364-372), S = support[p] (odd).                               */
typedef struct ddf_cf {
  const float*   data;      /* complex64, all planes packed                                     */
  const int64_t* offset;    /* [2*nw] complex-element offset of each plane inside data          */
  const int32_t* support;   /* [nw] support S of plane p (shared by Wplanes[p], WplanesConj[p]) */
  int32_t nw;               /* WInfos[2] = len(Wplanes)                                         */
  int32_t overs;            /* WInfos[3] = OverS                                                */
  double  ref_wave;         /* WInfos[0] = c / max(freq)                                        */
  double  wmax;             /* WInfos[1]                                                        */
  int32_t max_support;      /* host: max over support[] (sizes launch resources)                */
} ddf_cf;

/* The 20-entry LJones list (DDFacet/Imager/ClassDDEGridMachine.py:571-679, 759-776) as read by
 * DDEs::JonesServer (DDFacet/Gridder/JonesServer.cc:97-191).  enabled=0 <=> LJones == [].      */
typedef struct ddf_jones {
  int32_t enabled;
  int32_t mode;             /* LJones[9][1]: 1 = per visibility, 2 = per BDA block              */
  int32_t interp_mode;      /* LJones[9][0]: 0 = nearest, 1 = "Krigging"                        */
  const float*   kms;       /* LJones[0] complex64 [nt][nd][na][nch][2][2] or NULL              */
  int32_t kms_dims[4];      /* host: nt, nd, na, nch                                            */
  const int32_t* kms_time;  /* LJones[1]  [nrow]                                                */
  const int32_t* kms_chan;  /* LJones[10] [nchan]                                               */
  int32_t kms_dir;          /* LJones[6][0]                                                     */
  const float*   beam;      /* LJones[2] complex64 [nt][nd][na][nch][2][2] or NULL              */
  int32_t beam_dims[4];
  const int32_t* beam_time; /* LJones[3]  [nrow]                                                */
  const int32_t* beam_chan; /* LJones[11] [nchan]                                               */
  int32_t beam_dir;         /* LJones[8][0]                                                     */
  const int32_t* a0;        /* LJones[4] [nrow]                                                 */
  const int32_t* a1;        /* LJones[5] [nrow]                                                 */
  const float*   coefs;     /* LJones[7] [ncoefs] (interp_mode 1 only) or NULL                  */
  int32_t ncoefs;
  int32_t apply_amp;        /* LJones[13]                                                       */
  int32_t apply_phase;      /* LJones[14]                                                       */
  int32_t scale_jones;      /* LJones[15]                                                       */
  double  calib_error;      /* LJones[16]                                                       */
  double* sum_jones;        /* LJones[17] float64 [2][nband], accumulated (gridding only)       */
  double* sum_jones_chan;   /* LJones[18] float64 [2][nchan], accumulated (gridding only)       */
  double  reweight_snr;     /* LJones[19] (unused: every reference call site passes
                               EstimateWeight=false, gridder.h:236,294 degridder.h:220,245)    */
} ddf_jones;

/* The 7-entry LSmearing list (ClassDDEGridMachine.py:545-569, DecorrelationHelper.cc:25-41).   */
typedef struct ddf_decorr {
  int32_t enabled;          /* len(LSmearing) != 0                                              */
  const double* uvw_dt;     /* [nrow][3]                                                        */
  double  dt, dnu;
  int32_t tsmear, fsmear;
  double  l0, m0;
} ddf_decorr;

/* Visibility-side arguments shared by gridding and degridding.                                  */
typedef struct ddf_vis {
  float*         vis;       /* complex64 [nrow][nchan][ncorr] (read by grid, vis -= model by degrid) */
  const double*  uvw;       /* [nrow][3]                                                        */
  const uint8_t* flags;     /* bool [nrow][nchan][ncorr]                                        */
  const float*   weights;   /* [nrow][nchan] (gridding only)                                    */
  const double*  freqs;     /* [nchan]                                                          */
  const int32_t* chan_map;  /* [nchan] visibility channel -> grid band                          */
  int64_t nrow;
  int32_t nchan, ncorr;
  int32_t chan_equidistant; /* LOptimisation[1]                                                 */
  /* BDA block map, int32 [2 + nblocks + sum(2 + rows_b)] (DDFacet/Data/ClassSmearMapping.py:83-135) */
  const int32_t* bda;
  int64_t bda_len;          /* host: number of int32 words in bda                               */
  int64_t nblocks;          /* host: bda[0] + (bda[1] << 32)                                    */
  const int64_t* bda_start; /* optional [nblocks] word offset of block b's (chStart, chEnd, rows...)
                               record inside bda (from ddf_bda_index); NULL = computed per call */
  const uint8_t* sparsification; /* bool [nblocks] or NULL (gridding only)                      */
  uint16_t corr_codes[4];   /* DataCorrFormat                                                   */
  int32_t  ncorr_codes;
  uint16_t stokes_codes[4]; /* ExpectedOutStokes                                                */
  int32_t  nstokes;
  double   freq_mean;       /* host: mean of freqs[]      (0: computed from the device array)  */
  double   wavelength_mean; /* host: mean of c / freqs[]  (0: computed from the device array)  */
} ddf_vis;

/* One facet: its grid and its geometry.                                                         */
typedef struct ddf_facet {
  void*   grid;             /* complex64 (or complex128 if grid_f64) [nband][npol][ny][nx]      */
  double* sumwt;            /* float64 [nband][npol], accumulated (gridding only)               */
  ddf_cf  cf;
  double  cu, cv, l0, m0;   /* FacetInfos[0:4] (ClassDDEGridMachine.py:748-749)                  */
  ddf_jones jones;
  ddf_decorr decorr;
} ddf_facet;

typedef struct ddf_grid_desc {
  int32_t grid_f64;         /* 0: complex64 grid (pyGridderWPol32), 1: complex128 (pyGridderWPol64) */
  int32_t nband, npol, ny, nx;
  double  incr[2];          /* increment: (-cell, +cell) radians                                */
  int32_t dopsf;
} ddf_grid_desc;

/* ---- library ------------------------------------------------------------------------------- */
int         ddf_version(void);               /* DDF_ABI_VERSION                                  */
const char* ddf_last_error(void);
int         ddf_init(int device);            /* optional: select device, create workspace        */

/* Word offsets of every BDA block record, for reuse across calls (major cycles).
 * bda: device int32 map, out: device int64 [nblocks].                                           */
int ddf_bda_index(const int32_t* bda, int64_t bda_len, int64_t nblocks, int64_t* out, void* stream);

/* ---- plain single-facet entry points ---------------------------------------------------------
 * ddf_grid replaces pyGridderWPol<gridtype> (DDFacet/Gridder/GridderSmearPols.cc:76-209 ->
 * gridder::gridder, gridder.h:78-377): grid, sumwt, sum_jones, sum_jones_chan are accumulated.
 * ddf_degrid replaces pyDeGridderWPol (GridderSmearPols.cc:212-348 -> degridder::degridder,
 * degridder.h:55-266): vis -= model.  The caller passes the CF lists as the reference call sites
 * do: put() passes (Wplanes, WplanesConj), get() passes (WplanesConj, Wplanes)
 * (ClassDDEGridMachine.py:801-802, 1091-1092); here plane p < nw is the list given first.       */
int ddf_grid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, void* stream);
int ddf_degrid(const ddf_grid_desc* g, const ddf_vis* v, const ddf_facet* f, void* stream);

/* Weights scatter-add, replaces pyAccumulateWeightsOntoGrid[NoSem] (GridderSmearPols.cc:32-74):
 * grid[index[i]] += weights[i] for weights[i] != 0.  grid float64, weights float32, index int64. */
int ddf_accumulate_weights(double* grid, int64_t grid_len, const float* weights,
                           const int64_t* index, int64_t n, void* stream);

/* ---- imaging weights (uniform / Briggs) ------------------------------------------------------
 * The weighting steps of ClassVisServer._CalcWeights_serial for one chunk, on the device: cell
 * index, scatter-add (ddf_weights_accumulate), Briggs transform of the weight grid,
 * renormalisation.  Grid geometry as the reference sets it
 * (ClassVisServer.py:1122-1133): xymax = floor(uvmax / cell) + 1, npixx = 2 xymax + 1,
 * npix = npixx (xymax + 1), grid float64 [nbands][npix]. */

/* Replaces ClassVisServer._uv_to_index (ClassVisServer.py:983-1004): for row r and channel c,
 * (u, v) = uvw[r][0:2] negated when v < 0; x = floor(u freqs[c] / c0 / cell) + xymax,
 * y = floor(v freqs[c] / c0 / cell) (c0 = 299792458); index = y npixx + x, plus chan_band[c] npix
 * when nbands > 1, narrowed to int32 as numpy does; index = 0 where weights[r][c] == 0.
 * uvw float64 [nrow][3], weights float32 [nrow][nchan], freqs float64 [nchan], chan_band int32
 * [nchan] (unused when nbands == 1), index int32 [nrow][nchan] (output). */
int ddf_weights_index(const double* uvw, const float* weights, int64_t nrow, int nchan, const double* freqs,
                      const int32_t* chan_band, double cell, int64_t npix, int64_t npixx, int nbands,
                      int64_t xymax, int32_t* index, void* stream);

/* The scatter-add of _accumulateWeights_handler (ClassVisServer.py:1006-1019) for a [nrow][nchan]
 * weight array and its int32 index from ddf_weights_index: grid[index[r][c]] += weights[r][c] for
 * non-zero weights (pyAccumulateWeightsOntoGrid, GridderSmearPols.cc:32-74, without the widening
 * of the index to int64); indices outside [0, grid_len) are skipped. */
int ddf_weights_accumulate(double* grid, int64_t grid_len, const float* weights, const int32_t* index, int64_t nrow,
                           int nchan, void* stream);

/* Replaces the robust-weighting grid transform (ClassVisServer.py:1188-1197): per band,
 * avgW = sum g^2 / sum g and g <- 1 + g (5 10^-robust)^2 / avgW.  grid float64 [nbands][npix];
 * work: float64 scratch of ddf_weights_briggs_work_size(nbands) elements (per-workgroup partial
 * sums), work_len its length. */
int64_t ddf_weights_briggs_work_size(int nbands);
int ddf_weights_briggs(double* grid, int nbands, int64_t npix, double robust, double* work, int64_t work_len,
                       void* stream);
/* ddf_weights_briggs without the workspace argument, allocating nothing: the partial sums live in
 * a library-owned device array (256 KB per GPU, held until the process ends), so nbands <= 64.
 * No counterpart in the reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  Calls that overlap on one GPU share the array and race; results equal
 * ddf_weights_briggs's. */
int ddf_weights_briggs_noalloc(double* grid, int nbands, int64_t npix, double robust, void* stream);

/* Replaces the renormalisation of ClassVisServer._finalizeWeights_handler
 * (ClassVisServer.py:1035-1040): weights[i] /= grid[index[i]] with indices >= grid_len set to 0
 * (negative ones count from the end, as numpy indexing).  weights float32 (in place), index int32,
 * grid float64 [grid_len]. */
int ddf_weights_finalize(float* weights, const int32_t* index, int64_t n, const double* grid, int64_t grid_len,
                         void* stream);

/* ---- chunk flag preparation ------------------------------------------------------------------
 * Replaces the flag preparation of the chunk reader (ClassMS.ReadData, DDFacet/Data/ClassMS.py:
 * 1155-1167), which the gridders rely on (they test the flag of correlation 0 only): rows with
 * a0 == a1 flagged; visibilities with a NaN part flagged; every visibility flagged so far set to
 * 1e9 + 0i; then all correlations of a (row, channel) flagged when any is.  flags uint8 (0 / 1)
 * [nrow][nchan][ncorr] and data complex64 (interleaved re, im) [nrow][nchan][ncorr], both in place;
 * data may be null (flags only); a0, a1 int32 [nrow]; ncorr 1..4. */
int ddf_prepare_flags(uint8_t* flags, float* data, const int32_t* a0, const int32_t* a1, int64_t nrow, int nchan,
                      int ncorr, void* stream);

/* ---- batched major cycle: every facet of a chunk in a few persistent launches ----------------
 * Replaces the per-facet job loops of ClassFacetMachine (DDFacet/Imager/ClassFacetMachine.py:
 * 1574-1592 gridding, 1912-1943 degridding, one APP job per facet) and the row-semaphore
 * reduction of the degridder (degridder.h:236-263, Semaphores.cc:27-58).  No reference entry
 * point corresponds one to one: a caller that used to loop pyGridderWPol32 / pyDeGridderWPol over
 * the facets of a chunk builds one plan and calls ddf_plan_grid / ddf_plan_degrid per cycle.
 *
 * A plan depends on uvw, flags, frequencies, channel / BDA maps, CF geometry and facet geometry,
 * not on visibility or model values: build it once per chunk and reuse it every major cycle.
 * facets[i].cf lists the planes in put() order (Wplanes first); models[i] is facet i's complex64
 * model grid for degridding; facets[i].grid receives facet i's gridded visibilities.  All facets
 * share nw, overs, ref_wave, wmax and the per-plane supports (as ClassWTermModified builds them).
 * The Jones tables are taken from facets[0].jones; each facet supplies its direction indices and
 * its sum_jones / sum_jones_chan outputs.  vgrid / vdegrid carry the grid / degrid BDA maps.   */
typedef struct ddf_plan ddf_plan;

typedef struct ddf_plan_opts {
  int32_t tile;            /* uv tile edge in grid cells (0: 16)                                */
  int32_t chunk;           /* max records per tile work item (0: 2048)                          */
  int32_t blocks_per_item; /* BDA blocks per gridding work item (0: 128)                        */
  int32_t group;           /* facets per lane group in gridding (0: auto, <= 64)                */
  int32_t precision;       /* 0 mixed: FP64 uvw rotation and per-row phase base, FP32 sin/cos and
                              FP32 phasor advance along channels (re-anchored every 16), FP32
                              run sums / Jones products / Stokes maps, FP64 stamp, facet and
                              weight sums;
                              1 FP64 arithmetic throughout (as the reference);
                              2 FP32 throughout (phase formed and reduced in FP32)              */
  int32_t workgroups;      /* persistent workgroups per launch (0: 4 per compute unit for the
                              per-block passes, as many as fit on the device for the tile passes) */
} ddf_plan_opts;

int ddf_plan_create(const ddf_grid_desc* g, const ddf_vis* vgrid, const ddf_vis* vdegrid,
                    const ddf_facet* facets, void* const* models, int32_t nfacets,
                    const ddf_plan_opts* opts, ddf_plan** plan, void* stream);
/* Grid `vis` (NULL: vgrid->vis) onto every facet grid; accumulates grids, sumwt and the Jones
 * sums exactly like one pyGridderWPol32 call per facet.                                         */
int ddf_plan_grid(ddf_plan* plan, float* vis, int32_t dopsf, void* stream);
/* vis (NULL: vdegrid->vis) -= sum over facets of the degridded models, like one
 * pyDeGridderWPol call per facet.                                                                */
int ddf_plan_degrid(ddf_plan* plan, float* vis, void* stream);
/* Adjoint of ddf_plan_degrid as a linear map from the facet model grids to the visibilities:
 * facets[i].grid += D_i^H vis_bar for every facet (conjugate phasor, J0^H V J1 Jones, transposed
 * Stokes map, conjugate CF stamps through the degrid map's placements).  No reference
 * counterpart: it is the backward of the differentiable degridding operator.  The caller zeroes
 * the grids first; vis_bar is complex64 [nrow][nchan][ncorr].                                   */
int ddf_plan_degrid_adjoint(ddf_plan* plan, const float* vis_bar, void* stream);
/* Counters: [0] grid records, [1] degrid records, [2] grid work items, [3] degrid work items,
 * [4] degrid visibilities, [5] device bytes held, [6] tile, [7] facet group, [8]-[10] persistent
 * workgroups of the grid tile pass, the degrid tile pass and the adjoint tile pass.            */
int ddf_plan_info(const ddf_plan* plan, int64_t* out, int32_t n);
/* The plan's integer placement maps, a debug dump of what gridder.h:321-351 / degridder.h:170-200
 * compute per block (no reference entry point returns them; the reference's maps are observable
 * only through the cells a block touches): for the grid (degrid = 0) or degrid map, one 64-bit
 * record per (block, facet), block-major, computed on `stream` into the device buffer out (n >=
 * nblocks * nfacets): locx bits 0-15, locy 16-31, CF plane 32-39 (put order: Wplanes iw, then
 * WplanesConj nw + iw), oversampling sub-kernel io 40-47 and jo 48-55, band 56-62, valid bit 63
 * (0: block dropped, flagged, sparsified away or outside the band range).  The records are
 * recomputed from the chunk's arrays given to ddf_plan_create, which must still be valid (as for
 * every call on the plan).                                                                    */
int ddf_plan_placements(const ddf_plan* plan, int32_t degrid, uint64_t* out, int64_t n, void* stream);
/* Per-pass timing, a measurement hook with no reference counterpart: with timing on, every pass
 * of ddf_plan_grid / ddf_plan_degrid / ddf_plan_degrid_adjoint is bracketed by HIP events on the
 * launch stream; ddf_plan_pass_ms waits for the latest recording of each pass and writes its
 * duration in ms (-1 for a pass not yet run) for passes 0..n-1 in the order below.             */
#define DDF_PASS_GRID_A 0   /* grid pass A: per-block phased Jones-corrected block sums        */
#define DDF_PASS_GRID_B 1   /* grid pass B: LDS-tiled CF stamps into the facet grids           */
#define DDF_PASS_DEGRID_A 2 /* degrid pass A: LDS-staged CF gathers of the model grids          */
#define DDF_PASS_DEGRID_B 3 /* degrid pass B: per-visibility sum over facets, subtract        */
#define DDF_PASS_ADJ_B 4    /* adjoint of degrid pass B                                        */
#define DDF_PASS_ADJ_A 5    /* adjoint of degrid pass A                                        */
#define DDF_PLAN_NPASS 6
int ddf_plan_timing(ddf_plan* plan, int32_t enable);
int ddf_plan_pass_ms(ddf_plan* plan, float* ms, int32_t n);
int ddf_plan_destroy(ddf_plan* plan);

/* ---- chunk staging --------------------------------------------------------------------------
 * The baseline-time sort of ClassMS.ReadData (DDFacet/Data/ClassMS.py:575-586):
 * index[k] = the row of rank k in sorted(zip(A0, A1, TIME, range(nrow))) (antennas < 65536).
 * a0 / a1 int32, times f64, index int32: device arrays of nrow entries.                         */
int ddf_sort_baseline_time(const int32_t* a0, const int32_t* a1, const double* times, int64_t nrow, int32_t* index,
                           void* stream);
/* dst[k] = src[index[k]] for rows of row_bytes bytes (applies the sort index to uvw, data, flags,
 * weights, ... as ReadData does with sort_index); src and dst must not overlap.                  */
int ddf_gather_rows(const void* src, void* dst, int64_t nrow, int64_t row_bytes, const int32_t* index, void* stream);
/* The selections of ClassMS.UpdateFlags (ClassMS.py:1169-1233) that follow ddf_prepare_flags, on
 * flags uint8 [nrow][nchan][ncorr] equalised over the correlations, in place:
 *   per antenna the fraction of flagged (row, channel) cells of its rows (counted before the cuts
 *   below, as the reference counts flags1) -> ant_frac f64 [na] (device);
 *   rows with u^2 + v^2 outside [uv_min_m2, uv_max_m2] (use_uv_range; the reference's
 *   UVRangeKm d -> d^2 1e6), rows with times - time_offset outside [time_min, time_max]
 *   (use_time_range; TimeRange on unix times, time_offset = MJD-seconds to unix seconds) flagged;
 *   every row of an antenna with a fraction above ant_threshold (0.9 in the reference) or named
 *   in named_antennas (host uint8 [na]: DistMaxToCore, FlagAnts; may be NULL) flagged;
 *   ant_flagged uint8 [na] (device) receives the antennas flagged.                               */
typedef struct ddf_flag_select {
  int32_t na, use_uv_range, use_time_range, pad;
  double uv_min_m2, uv_max_m2;
  double time_min, time_max, time_offset;
  double ant_threshold;
  const uint8_t* named_antennas;
} ddf_flag_select;
int ddf_flag_selection(uint8_t* flags, const double* uvw, const double* times, const int32_t* a0, const int32_t* a1,
                       int64_t nrow, int32_t nchan, int32_t ncorr, const ddf_flag_select* sel, double* ant_frac,
                       uint8_t* ant_flagged, void* stream);

/* ---- BDA block maps built on the device ------------------------------------------------------
 * Replaces ClassSmearMapping.computeSmearMappingInBackground + collectSmearMapping
 * (DDFacet/Data/ClassSmearMapping.py:64-135): one job per baseline running
 * GiveBlocksRowsListBL_old (mode 1, :496-625, the default BDAMode) or the phase-based builder
 * (mode 2, :348-491, as restated in ddfacet_amd/bda.py), collected in lexicographic order of the
 * "a0:a1" keys into the int32 word format [NTot_lo, NTot_hi, sizes..., (chStart, chEnd, rows...)...].
 * The scalars are the builder's, formed on the host in its operation order:
 *   mode 1: duv_cut = C dPhi / (pi l nu0), k_dnu = (C / pi) dPhi, l, dfreq (C = 3e8);
 *   mode 2: k_phase = 2 pi nu0 / C, lmax, nmax = |sqrt(1 - lmax^2) - 1|, k_span = (C / (2 pi)) dPhi,
 *           dphi, dfreq, per_row = (C dPhi / (2 pi nu0) == 0).
 * uvw f64 [nrow][3], a0 / a1 int32 [nrow], chan_map int32 [nchan]: device pointers; rank: host
 * int32 [na][na], the position of "a0:a1" among the sorted keys of every ordered pair a0 != a1.
 * create sorts the rows by baseline and counts the map (it synchronises `stream` to return
 * nwords); write fills a device buffer of at least nwords int32 words.                          */
typedef struct ddf_bda_spec {
  int32_t mode, nchan, na, per_row;
  double duv_cut, k_dnu, l, dfreq;
  double k_phase, lmax, nmax, k_span, dphi;
} ddf_bda_spec;
typedef struct ddf_bda_builder ddf_bda_builder;
int ddf_bda_build_create(const ddf_bda_spec* spec, const double* uvw, const int32_t* a0, const int32_t* a1,
                         int64_t nrow, const int32_t* chan_map, const int32_t* rank, ddf_bda_builder** builder,
                         int64_t* nwords, void* stream);
int ddf_bda_build_write(ddf_bda_builder* builder, int32_t* out, int64_t n, void* stream);
int ddf_bda_build_destroy(ddf_bda_builder* builder);

/* ---- on-device CF tables ----------------------------------------------------------------------
 * Replaces the w-plane loop of ModCF.ClassWTermModified (DDFacet/Imager/ModCF.py:444-597) for
 * many facets at once: per facet f and plane i of support S_i the S x S screen
 * exp(-2 pi i (wlist_i / ref_wave) n1_f(l, m)) |taper_i|, zero-padded to (S_i overs)^2, centred
 * FFT / N^2 in FP64, complex64, reorganised to [overs][overs][S][S]; the conjugate planes from
 * the conjugate screen.  The host supplies what ModCF computes once per facet or support:
 * coefs[f][36] the order-5 polyval2d coefficients of n(l, m) - n0 with the linear terms zeroed
 * (ModCF.Give_dn, ModCF.py:267-295), tapers the |SpheMachine.MakeSphe(S_i)| of every plane
 * concatenated ([nw][S_i^2]), lrad the facet half-width in radians.
 * out (device, complex64): nfacets blocks of the 2*nw planes in put() order, back to back;
 * plane_offset (host, [2*nw]) receives each plane's offset in a facet block (in complex
 * elements), the same for every facet.                                                          */
typedef struct ddf_cf_spec {
  int32_t nw, overs, nfacets, pad_;
  const int32_t* sups;   /* host [nw], odd                                                      */
  const double* wlist;   /* host [nw], linspace(0, wmax, nw)                                     */
  const double* coefs;   /* host [nfacets][36]                                                   */
  const double* tapers;  /* host [sum of S_i^2]                                                  */
  double ref_wave, lrad;
} ddf_cf_spec;

int ddf_cf_build(const ddf_cf_spec* spec, float* out, int64_t* plane_offset, void* stream);

/* ---- facet image side of the major cycle ---------------------------------------------------
 * One facet of the image plane.  Device pointers; n = npix_facet (the padded facet size).      */
typedef struct ddf_image_facet {
  int32_t xc, yc;           /* DicoImager[i]["pixCentral"]: facet centre in main-image pixels    */
  const float* sphe;        /* float32 [n][n] cf_dict["Sphe"]                                    */
  const float* inv_sphe;    /* float32 [n][n] cf_dict["InvSphe"]                                 */
  const float* sw;          /* float32 [n][n] cf_dict["SW"] (the stored spatial weight)          */
  const float* dirty;       /* complex64 [nband][npol][n][n] facet image after GridToIm          */
  float* model;             /* complex64 [nband][npol][n][n] model grid written by model_to_grid */
  const double* sumwt;      /* float64 [nband][npol]                                             */
  const double* sum_jones;  /* float64 [2][nband], or NULL without Jones                         */
} ddf_image_facet;

/* ---- Beam Jones producers ------------------------------------------------------------------------
 * ddf_beam_jones: the analytic primary-beam E-Jones of ClassJones.EstimateBeam
 * (DDFacet/Data/ClassJones.py:1014-1065) for the GMRT and ATCA models
 * (ClassGMRTBeam.GiveRawBeam / GiveInstrumentBeam, DDFacet/Data/ClassGMRTBeam.py:140-173;
 * ClassATCABeam.GiveRawBeam / GiveInstrumentBeam, DDFacet/Data/ClassATCABeam.py:191-245): the beam
 * of each direction (ra, dec) from the phase centre (ra_c, dec_c), averaged over the channels of
 * each Jones channel, optionally normalised by the inverse of the beam at the phase centre
 * (CenterNorm), written as complex64 [ntime][ndir][na][nchan_jones][2][2] -- the layout of the
 * ddf_jones tables, so the output feeds the gridders directly.  The analytic models depend on
 * neither time nor antenna; ntime and na set the table's shape.  All pointers are device memory.
 * Per-channel model terms (ascending polynomial coefficients in float32-rounded values, null
 * thresholds) come from the host (ddfacet_amd/beams.py restates calcCoefs).                     */
enum { DDF_BEAM_GMRT = 1, DDF_BEAM_ATCA = 2 };
typedef struct ddf_beam_desc {
  int32_t kind;            /* DDF_BEAM_GMRT or DDF_BEAM_ATCA */
  int32_t nchan;           /* visibility channels */
  int32_t ncoef;           /* coefficients per polynomial, <= 16 */
  int32_t nchan_jones;
  const double* freqs;     /* [nchan] Hz */
  const double* coef0;     /* [nchan][ncoef]: GMRT the channel's band; ATCA the band at or below */
  const double* coef1;     /* [nchan][ncoef]: ATCA the band at or above (GMRT: unused) */
  const double* band_f0;   /* [nchan][2]: ATCA the two bands' frequencies (GMRT: unused) */
  const float* x_null;     /* [nchan]: GMRT arcmin; ATCA arcmin GHz */
  const float* y_null;     /* [nchan]: GMRT the clipped value (ATCA: computed here) */
  const int32_t* chan_to_jones; /* [nchan] */
} ddf_beam_desc;
int ddf_beam_jones(const ddf_beam_desc* b, const double* ra, const double* dec, int32_t ndir, double ra_c,
                   double dec_c, int32_t ntime, int32_t na, int32_t center_norm, void* jones, void* stream);

/* ddf_jones_merge: the cell products of ClassJonesDomains.MergeJones
 * (DDFacet/Other/ClassJonesDomains.py:142-212, behind ClassJones.MergeJones, ClassJones.py:1067-1072):
 * out[t][d][a][c] = J0[it0[t]][d][a][ich0[c]] . J1[it1[t]][d][a][ich1[c]] (2 x 2 complex64
 * products, ModLinAlg.BatchDot), the identity where ich0[c] or ich1[c] is -1.  The merged time and
 * frequency domains and the index maps (it*: nearest centre time, ich*: containing channel) are
 * small host arrays (ddfacet_amd/beams.py); all pointers are device memory.                      */
typedef struct ddf_jones_merge_desc {
  const void* j0;          /* complex64 [nt0][nd][na][nch0][2][2] */
  const void* j1;          /* complex64 [nt1][nd][na][nch1][2][2] */
  void* out;               /* complex64 [nt][nd][na][nch][2][2] */
  const int32_t* it0;      /* [nt] */
  const int32_t* it1;      /* [nt] */
  const int32_t* ich0;     /* [nch] */
  const int32_t* ich1;     /* [nch] */
  int32_t nt0, nch0, nt1, nch1, nd, na, nt, nch;
} ddf_jones_merge_desc;
int ddf_jones_merge(const ddf_jones_merge_desc* m, void* stream);

/* Centred 2-D FFT of nplanes n x n complex64 planes in place, ModFFTW.FFTW_2Donly
 * (DDFacet/ToolsDir/ModFFTW.py:557-600): forward fftshift(fft2(ifftshift(A))) / n^2, inverse
 * fftshift(ifft2(ifftshift(A))) * n^2, both times `scale`.  GridToIm
 * (ClassDDEGridMachine.py:1186-1190) is ddf_fft2c(grid, n, nband*npol, 1, OverS^2).            */
int ddf_fft2c(float* planes, int32_t n, int64_t nplanes, int32_t inverse, double scale, void* stream);
/* BuildFacetNormImage (ClassFacetMachine.py:1289-1325): facet_norm float32 [m][m].              */
int ddf_facet_norm(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t m,
                   float* facet_norm, void* stream);
/* FacetsToIm_Channel, kind Dirty / PSF (ClassFacetMachine.py:1331-1416): image float32
 * [nband][npol][m][m] (overwritten).                                                             */
int ddf_stitch(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband,
               int32_t npol, int32_t m, const float* facet_norm, float* image, void* stream);
/* GiveModelTessel(ToGrid=False, ApplyNorm=True) (ClassImToGrid.py:138-264) and the forward FFT of
 * get(TranformModelInput="FT") (ClassDDEGridMachine.py:951-958): writes facets[i].model from
 * the main model image float32 [nband][npol][m][m].                                               */
int ddf_model_to_grid(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband,
                      int32_t npol, int32_t m, const float* image, const float* facet_norm,
                      int32_t overs, void* stream);

/* Adjoint of ddf_model_to_grid (no reference counterpart; backward of the differentiable
 * forward model): takes facets[i].model as the model-grid gradient (overwritten by the adjoint
 * FFT) and writes image_bar float32 [nband][npol][m][m] (overwritten).                           */
int ddf_model_to_grid_adjoint(const ddf_image_facet* facets, int32_t nf, int32_t npix_facet, int32_t nband,
                              int32_t npol, int32_t m, float* image_bar, const float* facet_norm,
                              int32_t overs, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DDFGPU_H */
