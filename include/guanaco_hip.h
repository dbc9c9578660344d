/*
 * guanaco_hip.h - C ABI of the MI355X (gfx950) reconstruction library
 * libguanaco_hip.so.
 *
 * Plain C: pointers, sizes and status codes only. Every entry point returns
 * GX_OK (0) or a non-zero GX_E* code; gx_last_error() then describes the
 * failure (thread-local). All entry points are thread-safe; none but
 * gx_recon_noalloc uses device globals, so calls on different streams may
 * overlap (gx_recon_noalloc calls on one GPU may not).
 *
 * The reference's GPU path is reached through pybind11 (guanaco_ext.recon /
 * guanaco_ext.corr) and C++ classes; the entries below replace those classes
 * one for one (reference file:line given with each), and add a
 * device-resident whole-volume path that the reference does not have.
 */
#ifndef GUANACO_HIP_H
#define GUANACO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GX_OK 0
#define GX_EINVAL 1       /* invalid argument or configuration */
#define GX_ENOMEM 2       /* device allocation failed */
#define GX_EDEVICE 3      /* HIP runtime / kernel error */
#define GX_EUNSUPPORTED 4 /* size not supported by the gfx950 kernels */

/* Storage precision of filtered projections (accumulation is always FP32). */
#define GX_STORE_F32 0  /* exact reference FP32 arithmetic */
#define GX_STORE_F16 1  /* within 1e-3 rel-L2 of the FP64 reconstruction */
#define GX_STORE_BF16 2 /* 8-bit mantissa: about 2e-3, above that gate */

/* ---- status / devices ---------------------------------------------------- */

/* Message of the last failure on this thread (GUANACO_ASSERT's text in the
 * reference, include/guanaco/error.h:31-61). */
const char *gx_last_error(void);
int gx_version(void);
int gx_device_count(int *count);
int gx_set_device(int device);
int gx_get_device(int *device);
/* name (NUL-terminated, truncated to name_len), total HBM bytes, CU count */
int gx_device_info(int device, char *name, size_t name_len, size_t *total_mem,
                   int *num_cus);
int gx_mem_get_info(size_t *free_bytes, size_t *total_bytes);

/* ---- thin device memory / stream layer ----------------------------------- */
/* Streams and events are opaque HIP handles (NULL = the default stream). */

int gx_malloc(void **ptr, size_t bytes);
int gx_free(void *ptr);
int gx_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int gx_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int gx_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
/* Fill with a 32-bit pattern by kernel (safe under stream capture). */
int gx_fill_u32(void *dst, uint32_t value, size_t count, void *stream);
/* Plain element-wise copy kernel (also the byte-counter calibration). */
int gx_copy_u32(void *dst, const void *src, size_t count, void *stream);
int gx_stream_create(void **stream);
int gx_stream_destroy(void *stream);
int gx_stream_synchronize(void *stream);
int gx_device_synchronize(void);
int gx_event_create(void **event);
int gx_event_destroy(void *event);
int gx_event_record(void *event, void *stream);
int gx_event_synchronize(void *event);
int gx_event_elapsed_ms(void *start, void *stop, float *ms);
/* Stream capture into a graph and replay (hipStreamBeginCapture /
 * hipStreamEndCapture / hipGraphInstantiate / hipGraphLaunch). */
int gx_graph_capture_begin(void *stream);
int gx_graph_capture_end(void *stream, void **graph_exec);
int gx_graph_launch(void *graph_exec, void *stream);
int gx_graph_destroy(void *graph_exec);

/* ---- reference-equivalent entry points (host buffers, synchronous) -------- */

/* Mirrors guanaco::Config (include/guanaco/config.h:29-76). */
typedef struct gx_recon_cfg {
  size_t num_pixels;   /* N, detector columns */
  size_t num_angles;   /* T */
  size_t num_defocus;  /* ndf */
  size_t grid_height;  /* H (depth) */
  size_t grid_width;   /* W */
  float pixel_size;    /* A */
  float centre;        /* rotation centre, pixels */
  float min_defocus;   /* A, relative to the defocus at the tilt axis */
  float max_defocus;
  const float *angles; /* host, num_angles radians */
  int gpu_index;       /* device to use; -1 = current */
} gx_recon_cfg;

/* Per-tilt constants exactly as reconstructor.cpp:36-47 computes them:
 * coef[4t..4t+3] = {cos/d, sin/d, offset, scale} (float). Host only. */
int gx_tilt_geometry(const gx_recon_cfg *cfg, float *coef);

/* Replaces Reconstructor_t<e_device>::operator() (reconstructor.cu:317-354)
 * with the CPU path's semantics (reconstructor.cpp:92-115): filters a copy
 * of sino (num_defocus x num_angles x num_pixels, host), overwrites rec
 * (grid_height x grid_width, host) with the backprojection. */
int gx_recon(const gx_recon_cfg *cfg, const float *sino, float *rec);

/* No counterpart in the reference. gx_recon for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): d_sino (num_defocus x num_angles
 * x num_pixels) and d_rec (grid_height x grid_width) are device buffers,
 * d_rec is overwritten with gx_recon's result (the same kernels, bit for
 * bit), d_sino is not modified. The filtered copy and the per-call tables
 * live in a 64 MiB device-global scratch array per GPU (no hipMalloc or
 * hipFree): sizes with num_defocus * num_angles * num_pixels * 4 bytes plus
 * tables above 64 MiB are refused with GX_EINVAL. Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU race.
 * Synchronizes `stream` (NULL: the default stream) on entry and return. */
int gx_recon_noalloc(const gx_recon_cfg *cfg, const float *d_sino,
                     float *d_rec, void *stream);

/* Backprojection only (reconstructor.cpp:14-90) of an already filtered
 * sinogram; rec is overwritten (the reference zeroes it first). */
int gx_project(const gx_recon_cfg *cfg, const float *filtered, float *rec);

/* Ramp filter only (filter.cpp:61-97 / filter.cu:42-100): every row of
 * rows (nrows x num_pixels, host) filtered into out. */
int gx_filter(const float *rows, float *out, size_t nrows, size_t num_pixels,
              int gpu_index);

/* Replaces Corrector<e_device, float>::correct (correct.cu:30-86) behind
 * guanaco::correct (correct.h:77-95): image (ysize x xsize), ctf complex64
 * interleaved (num_ctf x ysize x xsize), rec (num_ctf x ysize x xsize)
 * overwritten with Re(IFFT2_unnorm(FFT2(image) * -sign(Im ctf))).
 * Any xsize and ysize in [1, 8192]: powers of two (xsize >= 8, ysize >= 4)
 * take the power-of-two transforms, other sizes Bluestein transforms (of up
 * to 16384 points). */
int gx_corr(const float *image, const float *ctf, float *rec, size_t xsize,
            size_t ysize, size_t num_ctf, int gpu_index);

/* CTF images on the FFT-ordered grid, computed on the device in FP64 and
 * stored as complex64: out[k][y][x] (interleaved re, im) for the absolute
 * defocus defocus[k]. With envelope, CTF.get_ctf(w, h, ps) of the reference
 * (ctf.h:466-478: A exp(log Es + log Et) exp(-i (chi - phi)), envelopes of
 * ctf.h:188-223); without, CTF.get_ctf_simple(w, h, ps) (ctf.h:489-503).
 * recentre moves the zero frequency to (height/2, width/2) (numpy fftshift),
 * as generate_ctf(recentre=True) (plot.py:280-281). `out` is host memory of
 * num_defocus * height * width * 2 floats. Synchronous. */
typedef struct gx_ctf_desc {
  size_t width, height, num_defocus;
  double pixel_size;     /* A */
  const double *defocus; /* host, num_defocus, A */
  double wavelength;     /* A */
  double Cs;             /* A */
  double Ca;             /* A */
  double Pa;             /* rad */
  double dd;             /* defocus spread, A */
  double theta_c;        /* source spread, rad */
  double phi;            /* rad */
  int envelope;          /* 1: with envelopes */
  int recentre;          /* 1: zero frequency at the centre */
  int gpu_index;         /* -1 = current device */
} gx_ctf_desc;
int gx_ctf_grid(const gx_ctf_desc *desc, float *out);

/* ---- device-resident whole-volume path ------------------------------------ */

/* Input layouts for gx_volume_run. */
#define GX_INPUT_PROJECTIONS 0 /* (T, Y, X) raw projections */
#define GX_INPUT_CORRECTED 1   /* (T, ndf, Y, X) already CTF-corrected */
#define GX_INPUT_SINOGRAMS 2   /* (Y, ndf, T, X) sinogram order */

typedef struct gx_volume_desc {
  size_t num_pixels;   /* X = N */
  size_t num_rows;     /* Y, detector rows = volume slices */
  size_t num_angles;   /* T */
  size_t num_defocus;  /* ndf */
  size_t depth;        /* Z, grid height (central rows of the square grid) */
  size_t width;        /* W, grid width */
  float pixel_size;    /* A */
  float centre;        /* rotation centre, pixels (same for every row) */
  float min_defocus;   /* A, relative, as reconstruct() passes to recon() */
  float max_defocus;
  const float *angles;  /* host, T radians */
  const float *weights; /* host, T per-tilt multipliers, or NULL for ones */
  int storage;          /* GX_STORE_* */
  int input_layout;     /* GX_INPUT_* */
  /* On-device CTF correction of GX_INPUT_PROJECTIONS: when defocus is not
   * NULL, slab k is phase-flipped with the absolute defocus defocus[k] (A),
   * as correct_projections does (correct.py:217-318). */
  const double *defocus;
  double wavelength; /* A */
  double Cs;         /* A */
  double Ca;         /* A */
  double Pa;         /* rad */
  double phi;        /* rad */
  size_t tilt_chunk; /* tilts per correction launch; 0 = automatic */
  int gpu_index;     /* -1 = current device */
  /* Per-row rotation centres (host, num_rows pixels), as reconstruct()
   * accepts them (reconstruct.py:32-52, one per slice in mp.py:144-154), or
   * NULL for `centre` on every row. With per-row centres the plan
   * backprojects with the direct kernel, taps gathered per row. */
  const float *centres;
  /* Detector rows whose filtered slabs the plan holds at a time (rounded up
   * to a multiple of 32), or 0 for every row. A banded plan runs the path
   * band by band (gx_volume_run): each band is filtered (for CTF-corrected
   * input, the whole 2D correction again, its row pass only for the band)
   * and backprojected into its rows of the volume, so FP32 slabs of volumes
   * whose slabs do not fit beside them fit. No reference counterpart: the
   * reference filters and backprojects one slice at a time (mp.py:106-154),
   * and slice y needs row y of every slab only (reconstructor.cpp:50,72). */
  size_t band_rows;
} gx_volume_desc;

typedef struct gx_volume_plan gx_volume_plan;

/* Device bytes the plan holds (filtered slabs + workspace), excluding the
 * caller's input and volume: the admission footprint. */
int gx_volume_plan_bytes(const gx_volume_desc *desc, size_t *bytes);
int gx_volume_plan_create(const gx_volume_desc *desc, gx_volume_plan **plan);
int gx_volume_plan_destroy(gx_volume_plan *plan);
/* Enqueue the whole path on `stream`: d_in (layout per input_layout, device)
 * -> d_vol (Y x Z x W float32, device, overwritten). No allocation, no
 * memset: the call can be captured into a graph. Calls that share a plan
 * must not overlap. */
int gx_volume_run(gx_volume_plan *plan, const float *d_in, float *d_vol,
                  void *stream);
/* Stages of gx_volume_run, for timing: filter/correct into the plan's slabs,
 * and backprojection of the slabs (plans that hold every row). */
int gx_volume_filter(gx_volume_plan *plan, const float *d_in, void *stream);
int gx_volume_backproject(gx_volume_plan *plan, float *d_vol, void *stream);
/* Filter/correct detector rows [y0, y1) only into the plan's slabs (y0 a
 * multiple of the plan's rows per block, y1 - y0 at most its band rows);
 * d_in holds every row. Later backprojections take rows of this band. */
int gx_volume_filter_rows(gx_volume_plan *plan, const float *d_in, size_t y0,
                          size_t y1, void *stream);
/* Rows held per band (the plan's rows when unbanded), and the band the
 * slabs hold after the last filter pass. */
int gx_volume_plan_band(gx_volume_plan *plan, size_t *band_rows, size_t *y0,
                        size_t *y1);
/* Backprojection of detector rows [y0, y1) only, into d_vol holding those
 * rows ((y1 - y0) x Z x W). y0 must be a multiple of the plan's rows per
 * block, y1 too unless it is num_rows, and the rows must lie in the band
 * last filtered: volumes larger than the free HBM are written band by
 * band. */
int gx_volume_backproject_rows(gx_volume_plan *plan, size_t y0, size_t y1,
                               float *d_vol, void *stream);
int gx_volume_rows_per_block(gx_volume_plan *plan, size_t *rows);
/* Backprojection kernel of a plan (no reference counterpart; the reference
 * has one CPU and one texture-based GPU kernel). GX_BP_STAGED, the default,
 * copies each voxel tile's detector window into LDS per batch of tilts and
 * holds 1, 2 or 4 defocus slabs per tilt; when a tile spans more slabs
 * (narrow slabs, large |dscale|) plans use GX_BP_DIRECT, which gathers taps
 * straight from global memory. Both give identical results. */
#define GX_BP_DIRECT 0
#define GX_BP_STAGED 1
int gx_volume_plan_set_backprojector(gx_volume_plan *plan, int kind);
int gx_volume_plan_get_backprojector(gx_volume_plan *plan, int *kind,
                                     int *slabs_per_tilt);
/* The plan's filtered slabs: [k][t][y/YB][x][y%YB], element size by storage.
 * Plans with several defocus slabs compute, at each tilt t, only the slabs
 * [first[t], last[t]] that some voxel of the volume reads (no reference
 * counterpart: the reference corrects every slab); the others, and the rows
 * of the last row block past the plan's rows, hold whatever the buffer held
 * (the buffer is not cleared). */
int gx_volume_slabs(gx_volume_plan *plan, void **d_slab, size_t *bytes,
                    int *rows_per_block);
/* first[t], last[t] for every tilt (host arrays of num_angles ints; 0 and 0
 * with one slab). Synchronous. */
int gx_volume_plan_slab_range(gx_volume_plan *plan, int *first, int *last);

/* ---- autotuning (no reference counterpart) --------------------------------
 * The backprojection tile of a plan shape is chosen by measurement. A tile
 * is 32 x 8 voxels in (x, z) by `rows` detector rows in y, one voxel column
 * per thread; `rows` is also the row blocking of the filtered slabs. The
 * candidates are rows 8 and 16 for float32 slabs, 16 and 32 for float16 and
 * bfloat16 ones, each with the staged and the direct backprojector (staged
 * only where a tile spans at most 4 defocus slabs). gx_volume_plan_autotune
 * filters d_in once per row count, times that filter pass and every
 * backprojector on the plan's own slabs with HIP events on `stream`, keeps
 * the candidate with the least filter + backprojection time, and leaves
 * d_vol reconstructed with it. Every candidate computes the same volume.
 * The choice is remembered per (device model, plan shape) for the life of
 * the process, and plans of that shape created later start with it. The
 * call synchronizes `stream`, cannot be captured into a graph, and must not
 * overlap other calls on the same plan. */
#define GX_TUNE_MAX 8
typedef struct gx_tune_entry {
  int kind;         /* GX_BP_* */
  int rows;         /* rows per tile */
  float filter_ms;  /* filter pass with this row blocking */
  float bp_ms;      /* backprojection */
} gx_tune_entry;
typedef struct gx_tune_result {
  int kind, rows;   /* the choice */
  int n;            /* candidates timed */
  gx_tune_entry cand[GX_TUNE_MAX];
} gx_tune_result;
int gx_volume_plan_autotune(gx_volume_plan *plan, const float *d_in,
                            float *d_vol, void *stream, gx_tune_result *out);
/* The plan's current tile (rows per tile and backprojector), and setting
 * it to one of the candidates by hand; a new row count takes effect with
 * the next filter pass (gx_volume_run, gx_volume_filter). */
int gx_volume_plan_get_tile(gx_volume_plan *plan, int *rows, int *kind);
int gx_volume_plan_set_tile(gx_volume_plan *plan, int rows, int kind);
/* The remembered choices: written as text, one shape per line, read back
 * (merged into the current ones), or forgotten. */
int gx_tune_save(const char *path);
int gx_tune_load(const char *path);
int gx_tune_clear(void);

/* ---- seeded synthetic workloads (no reference counterpart) --------------- */

/* Tilt series of a random-ellipsoid phantom projected analytically in the
 * reconstruction geometry, optionally multiplied by the CTF sin(chi - phi)
 * in Fourier space, with Gaussian or Poisson noise; every value depends on
 * `seed` only. Writes (num_angles x num_rows x num_pixels) float32 to d_out
 * (device). Synchronous. */
typedef struct gx_synth_desc {
  size_t num_pixels;     /* X */
  size_t num_rows;       /* Y */
  size_t num_angles;     /* T */
  size_t depth;          /* Z of the phantom box (0 = X) */
  size_t width;          /* W of the phantom box (0 = X) */
  float centre;          /* rotation centre, pixels */
  const float *angles;   /* host, T radians */
  size_t num_ellipsoids; /* <= 512 */
  uint64_t seed;
  float semi_min, semi_max;       /* semi-axes, pixels */
  float density_min, density_max; /* densities */
  int noise;         /* 0 none; 1 Gaussian, sigma = noise_level * max;
                        2 Poisson, mean = noise_level * (1 + contrast *
                        signal / max|signal|) */
  float noise_level;
  float contrast;
  int ctf;           /* apply the CTF (X, Y powers of two) */
  double wavelength; /* A */
  double Cs;         /* A */
  double defocus;    /* A */
  double phi;        /* rad */
  double pixel_size; /* A */
} gx_synth_desc;

int gx_synth_projections(const gx_synth_desc *desc, float *d_out,
                         void *stream);

/* ---- serving (no reference counterpart) ---------------------------------
 * A request loop for whole-volume reconstructions of different sizes: the
 * server owns one device allocation (its budget) and admits requests in
 * arrival order as their footprint (gx_server_request_bytes: plan, input
 * copy, volume) fits and one of its streams is free; with backfill a later
 * request that fits may start while an earlier one waits. Requests run
 * concurrently, one stream each, with no allocation or device-wide
 * synchronisation; each result volume stays in server memory until
 * released. Results do not depend on what else is running. */
typedef struct gx_server gx_server;

typedef struct gx_server_cfg {
  size_t budget_bytes; /* device memory the server holds; 0 = 80% of free */
  int max_inflight;    /* requests running at once; 0 = 4 */
  int backfill;        /* 1: later requests that fit may start first */
  int gpu_index;       /* -1 = current device */
} gx_server_cfg;

#define GX_REQ_QUEUED 0
#define GX_REQ_RUNNING 1
#define GX_REQ_DONE 2
#define GX_REQ_FAILED 3
#define GX_REQ_CANCELLED 4

typedef struct gx_request_info {
  int state;               /* GX_REQ_* */
  double submit_s;         /* server clock (s since creation) */
  double start_s, end_s;   /* admitted, finished */
  double device_ms;        /* correction + filter + backprojection on device */
  double updates;          /* X * Y * Z * T */
  size_t footprint_bytes;  /* held while running */
  const float *volume;     /* device, (Y, Z, W) float32, until released */
  char error[256];         /* failure message */
} gx_request_info;

typedef struct gx_server_stats {
  size_t submitted, completed, failed, cancelled, queued, running;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submission to completion */
  double updates_per_s;   /* completed updates / (last end - first submit) */
  double span_s;
  size_t budget_bytes, used_bytes, peak_bytes;
} gx_server_stats;

int gx_server_create(const gx_server_cfg *cfg, gx_server **server);
/* Waits for the running requests; requests still queued are cancelled. */
int gx_server_destroy(gx_server *server);
/* Device bytes a request holds: the plan and, for host input, its copy
 * (while running), plus the volume (until released). */
int gx_server_request_bytes(const gx_volume_desc *desc, int input_on_host,
                            size_t *bytes);
/* The descriptor and its arrays are copied; `input` (layout per
 * desc->input_layout; host or device memory) must stay valid until the
 * request has finished. Refused (GX_ENOMEM) when the footprint exceeds the
 * budget. */
int gx_server_submit(gx_server *server, const gx_volume_desc *desc,
                     const float *input, int input_on_host, uint64_t *id);
/* Blocks until the request has finished (or timeout_s > 0 passes). */
int gx_server_wait(gx_server *server, uint64_t id, double timeout_s,
                   gx_request_info *info);
int gx_server_info(gx_server *server, uint64_t id, gx_request_info *info);
/* A queued request is dropped; a running one is reported cancelled when it
 * ends. */
int gx_server_cancel(gx_server *server, uint64_t id);
/* Frees a finished request's volume; the id is then unknown. */
int gx_server_release(gx_server *server, uint64_t id);
int gx_server_get_stats(gx_server *server, gx_server_stats *stats);

#ifdef __cplusplus
}
#endif

#endif
