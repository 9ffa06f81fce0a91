/*
 * flam3_render on the GPU: the host half of render_rectangle (rect.c:529-1286).
 *
 * The host keeps everything that is per genome or per temporal sample and
 * cheap - interpolation, precalc flags, selection tables, camera, palette,
 * filter banks, k1/k2 - and sends each temporal sample to the GPU as one
 * packed job.  The histogram and accumulator stay in HBM for the whole
 * render; only the final image comes back.  Validation, error returns,
 * ISAAC consumption of f->rc, stats and progress/abort behave as in the
 * reference.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "fl3_device.h"
#include "fl3_host.h"
#include "fl3_pack.h"

#define WHITE_LEVEL 255.0
#define PREFILTER_WHITE 255
#define FUSE_NORMAL 15   /* rect.c:36 */
#define FUSE_WALKER 100
#define FUSE_EARLYCLIP 100

/* Host threads that prepare requests or temporal samples (each one's work
 * is independent: its own frame or genome and camera). */
#define FL3_PREP_THREADS 8
/* Job records carry the job index in 16 bits. */
#define MAX_BATCH_JOBS 4096
/* Renders with a linear palette below this many plotted samples share
 * walker launches (with per-sample atomics). */
#define FL3_SHARED_MAX_SAMPLES (1LL << 24)
/* Histogram cells of one shared launch: its renders' joint cell space in at
 * most 8192 tiles of 4096 cells. */
#define FL3_SHARED_MAX_CELLS (32LL * 1000 * 1000)
/* ... and of one whose renders share a palette: 4-byte records hold a
 * 24-bit cell index. */
#define FL3_SHARED_NARROW_CELLS ((1LL << 24) - 2)

/* One launch over the *n temporal samples collected so far (nsamples[t]
 * plotted samples and keys[2t], keys[2t+1] for the t-th); frees their jobs
 * and empties the set.  Interpolated sample densities differ in their last
 * bits between temporal samples (rect.c:797, 830-833), so the counts may
 * differ by a few samples: one launch takes them all, its walkers sized by
 * the largest. */
static int flush_batch(void **jobs, size_t *bytes, int *n, const int64_t *nsamples, const uint64_t *keys, int exact,
                       const fl3h_render_opts *opts, double *bad, int *sat) {
   fl3d_iter_params ip;
   int rc, t, same = 1;
   int64_t nmax = nsamples[0];
   for (t = 1; t < *n; t++) {
      if (nsamples[t] != nsamples[0]) same = 0;
      if (nsamples[t] > nmax) nmax = nsamples[t];
   }
   memset(&ip, 0, sizeof(ip));
   /* The reference fuses 15 steps (100 with earlyclip) before each
    * 10000-sample sub-batch; a production walker fuses once before a few
    * hundred to a few thousand samples, so it fuses 100 steps to keep the
    * start transient out of the image. */
   ip.fuse = FUSE_WALKER;
   ip.nsamples = nmax;
   ip.job_nsamples = same ? NULL : nsamples;
   ip.key[0] = keys[0];
   ip.key[1] = keys[1];
   ip.exact = exact;
   ip.nwalkers = opts ? opts->nwalkers : 0;
   ip.deposit = opts ? opts->deposit : 0;
   ip.chunk = opts ? opts->chunk_samples : 0;
   rc = fl3d_iterate_batch((const void *const *)jobs, bytes, *n, &ip, keys, bad, sat);
   for (t = 0; t < *n; t++) {
      free(jobs[t]);
      jobs[t] = NULL;
   }
   *n = 0;
   return rc;
}

/* Temporal samples per walker launch. */
static int max_batch_jobs(const fl3h_render_opts *opts) {
   if (opts && opts->temporal_batch > 0 && opts->temporal_batch < MAX_BATCH_JOBS) return opts->temporal_batch;
   return MAX_BATCH_JOBS;
}

static int env_int(const char *name, int dflt) {
   const char *v = getenv(name);
   return v ? atoi(v) : dflt;
}

double flam3_render_memory_required(flam3_frame *f) {
   const flam3_genome *g = f->genomes;
   int bits = f->bits == 33 ? 32 : f->bits;
   int bytes = bits / 8;
   return (double)g[0].spatial_oversample * g[0].spatial_oversample * (double)g[0].width * g[0].height *
          bytes * 9.0;
}

/* Progress callback handling (rect.c:324-391): 2 pauses, 1 aborts. */
static int poll_progress(flam3_frame *f, double pct, int stage, double eta) {
   int rv;
   if (!f->progress) return 0;
   rv = f->progress(f->progress_parameter, pct, stage, eta);
   while (rv == 2) {
      usleep(100000);
      rv = f->progress(f->progress_parameter, pct, stage, eta);
   }
   return rv == 1;
}

/* Per-thread reseeding of rect.c:858-865: 16 words from the frame stream,
 * then ISAAC seeding.  Returns the device form (36 words). */
static void thread_state(randctx *frame_rc, uint32_t *s) {
   randctx t;
   int k;
   memset(&t, 0, sizeof(t));
   for (k = 0; k < FLAM3_RANDSIZ; k++) t.randrsl[k] = fl3_irand(frame_rc);
   fl3_isaac_init(&t, 1);
   s[0] = (uint32_t)t.randcnt;
   for (k = 0; k < 16; k++) s[1 + k] = (uint32_t)t.randrsl[k];
   for (k = 0; k < 16; k++) s[17 + k] = (uint32_t)t.randmem[k];
   s[33] = (uint32_t)t.randa;
   s[34] = (uint32_t)t.randb;
   s[35] = (uint32_t)t.randc;
}

typedef struct {
   int image_width, image_height, out_width;
   int gutter, de_offset, width, height;
   double ppux, ppuy, sample_density;
} geometry;

/* Camera of one temporal sample (rect.c:784-827). */
static int setup_camera(const flam3_frame *spec, const flam3_genome *cp, int field, int oversample,
                        geometry *geo, fl3h_camera *cam) {
   double t0, t1, shift = 0.0, corner0, corner1, scale, size0, size1, ang;
   if (cp->sample_density <= 0.0) {
      fprintf(stderr, "sample density (quality) must be greater than zero, not %g.\n", cp->sample_density);
      return 1;
   }
   scale = pow(2.0, cp->zoom);
   geo->sample_density = cp->sample_density * scale * scale;
   geo->ppux = cp->pixels_per_unit * scale;
   geo->ppuy = field ? (geo->ppux / 2.0) : geo->ppux;
   geo->ppux /= spec->pixel_aspect_ratio;
   switch (field) {
      case flam3_field_even: shift = -0.5; break;
      case flam3_field_odd: shift = 0.5; break;
      default: shift = 0.0; break;
   }
   shift = shift / geo->ppux;
   t0 = (double)geo->gutter / (oversample * geo->ppux);
   t1 = (double)geo->gutter / (oversample * geo->ppuy);
   corner0 = cp->center[0] - geo->image_width / geo->ppux / 2.0;
   corner1 = cp->center[1] - geo->image_height / geo->ppuy / 2.0;
   cam->bounds[0] = corner0 - t0;
   cam->bounds[1] = corner1 - t1 + shift;
   cam->bounds[2] = corner0 + geo->image_width / geo->ppux + t0;
   cam->bounds[3] = corner1 + geo->image_height / geo->ppuy + t1 + shift;
   size0 = 1.0 / (cam->bounds[2] - cam->bounds[0]);
   size1 = 1.0 / (cam->bounds[3] - cam->bounds[1]);
   ang = cp->rotate * 2 * M_PI / 360.0;
   cam->rot[0][0] = cos(ang);
   cam->rot[0][1] = -sin(ang);
   cam->rot[1][0] = -cam->rot[0][1];
   cam->rot[1][1] = cam->rot[0][0];
   cam->rot_center[0] = cp->rot_center[0];
   cam->rot_center[1] = cp->rot_center[1];
   cam->rotate = cp->rotate != 0.0;
   cam->width = geo->width;
   cam->height = geo->height;
   cam->ws0 = geo->width * size0;
   cam->wb0s0 = cam->ws0 * cam->bounds[0];
   cam->hs1 = geo->height * size1;
   cam->hb1s1 = cam->hs1 * cam->bounds[1];
   return 0;
}

/* ---------------------------------------------------------------- render plan
 *
 * Everything of one render that the host derives: validation, interpolated
 * genome, geometry, filters, walker keys and the per-temporal-sample jobs.
 * fl3h_render runs one plan at a time; fl3h_render_many shares the walker
 * launch of several plans. */
typedef struct {
   flam3_frame *spec;
   void *out;                     /* first output row (odd fields start one row down) */
   int field, nchan, transp, bpc;
   stat_struct *stats;
   const fl3h_render_opts *opts;
   int replay, nthreads, cmap_size, oversample, nbatches, nts, filter_width;
   flam3_genome cp;
   geometry geo;
   fl3h_camera *cam;
   double *filter, *temporal_filter, *temporal_deltas, sumfilt;
   double vibrancy, gamma, background[3], badvals;
   int vib_gam_n;
   uint64_t *keys;                /* two per temporal sample */
   uint32_t *walker_states;       /* replay: 36 words per reference thread */
   time_t tstart;
} render_plan;

static void plan_close(render_plan *P) {
   free(P->filter);
   free(P->temporal_filter);
   free(P->temporal_deltas);
   free(P->cam);
   free(P->walker_states);
   free(P->keys);
   fl3_genome_free(&P->cp);
   memset(P, 0, sizeof(*P));
}

/* Validation, geometry and filters of rect.c:529-701.  Returns 1 after a
 * message on stderr for invalid parameters. */
static int plan_open(render_plan *P, flam3_frame *spec, void *out, int field, int nchan, int transp,
                     stat_struct *stats, const fl3h_render_opts *opts) {
   int i, max_de_fw;
   geometry *geo = &P->geo;
   memset(P, 0, sizeof(*P));
   P->tstart = time(NULL);
   P->spec = spec;
   P->field = field;
   P->nchan = nchan;
   P->transp = transp;
   P->bpc = spec->bytes_per_channel;
   P->stats = stats;
   P->opts = opts;
   P->replay = opts && opts->replay;
   if (spec->bits != 32 && spec->bits != 33 && spec->bits != 64) {
      fprintf(stderr, "flam3: bits must be 32, 33, or 64 not %d.\n", spec->bits);
      return 1;
   }
   P->nthreads = spec->nthreads > 0 ? spec->nthreads : 1;
   stats->num_iters = 0;
   /* apophysis palettes use 255 entries (rect.c:587) */
   P->cmap_size = 256 - env_int("apo_palette", 0);

   if (fl3_interpolate(spec->genomes, spec->ngenomes, spec->time, 0, &P->cp)) goto fail;
   P->oversample = P->cp.spatial_oversample;
   P->nbatches = P->cp.nbatches;
   P->nts = P->cp.ntemporal_samples;
   if (P->nbatches < 1) {
      fprintf(stderr, "nbatches must be positive, not %d.\n", P->nbatches);
      goto fail;
   }
   if (P->oversample < 1) {
      fprintf(stderr, "oversample must be positive, not %d.\n", P->oversample);
      goto fail;
   }

   geo->image_width = P->cp.width;
   geo->out_width = P->cp.width;
   if (field) {
      geo->image_height = P->cp.height / 2;
      if (field == flam3_field_odd) out = (unsigned char *)out + nchan * P->bpc * geo->out_width;
      geo->out_width *= 2;
   } else
      geo->image_height = P->cp.height;
   P->out = out;

   P->filter_width = fl3_create_spatial_filter(&spec->genomes[0], spec->pixel_aspect_ratio, field, &P->filter);
   if (P->filter_width < 0) {
      fprintf(stderr, "flam3_create_spatial_filter returned error: aborting\n");
      goto fail;
   }
   P->sumfilt = fl3_create_temporal_filter(P->nbatches * P->nts, P->cp.temporal_filter_type,
                                           P->cp.temporal_filter_exp, P->cp.temporal_filter_width,
                                           &P->temporal_filter, &P->temporal_deltas);

   /* gutter for the spatial filter, widened for the DE radius (rect.c:652-681) */
   geo->gutter = (P->filter_width - P->oversample) / 2;
   max_de_fw = 0;
   for (i = 0; i < spec->ngenomes; i++) {
      int w = (int)ceil(spec->genomes[i].estimator) * P->oversample;
      if (w > max_de_fw) max_de_fw = w;
   }
   if (max_de_fw > 0) max_de_fw += P->oversample - 1;
   if (max_de_fw > geo->gutter) {
      geo->de_offset = max_de_fw - geo->gutter;
      geo->gutter = max_de_fw;
   } else
      geo->de_offset = 0;
   geo->height = P->oversample * geo->image_height + 2 * geo->gutter;
   geo->width = P->oversample * geo->image_width + 2 * geo->gutter;

   P->cam = (fl3h_camera *)calloc(1, sizeof(fl3h_camera));
   P->walker_states = (uint32_t *)malloc(sizeof(uint32_t) * 36 * (size_t)P->nthreads);
   P->keys = (uint64_t *)calloc((size_t)2 * P->nts, sizeof(uint64_t));
   if (!P->cam || !P->walker_states || !P->keys) goto fail;
   return 0;
fail:
   plan_close(P);
   return 1;
}

/* The batch's density-estimation filters, from the genome at its first
 * temporal sample (rect.c:726-747). */
static int plan_batch_de(render_plan *P, int batch_num, fl3_de_filters *de) {
   memset(de, 0, sizeof(*de));
   if (fl3_interpolate(P->spec->genomes, P->spec->ngenomes,
                       P->spec->time + P->temporal_deltas[batch_num * P->nts], 0, &P->cp))
      return 1;
   if (P->cp.estimator < 0.0 || P->cp.estimator_minimum < 0.0) {
      fprintf(stderr, "density estimator filter widths must be >= 0\n");
      return 1;
   }
   if (P->spec->bits <= 32 && P->cp.estimator > 0.0) {
      fprintf(stderr, "warning: density estimation disabled with %d bit buffers.\n", P->spec->bits);
      P->cp.estimator = 0.0;
   }
   if (P->cp.estimator > 0.0) {
      if (fl3_create_de_filters(P->cp.estimator, P->cp.estimator_minimum, P->cp.estimator_curve, P->oversample,
                                de)) {
         fprintf(stderr, "de.kernel_size returned 0 - aborting.\n");
         return 1;
      }
   } else
      de->max_filter_index = 0;
   return 0;
}

/* Genome, tables, palette and camera of temporal sample tidx packed as one
 * job (rect.c:754-835).  first: the first pass over the batch (counts
 * num_iters and the tone-map averages). */
static int plan_ts_job(render_plan *P, int tidx, int first, void **job, size_t *job_bytes, double *batch_size) {
   flam3_genome *cp = &P->cp;
   const double color_scalar = P->temporal_filter[tidx];
   int *starts;
   int i, j, nrows, ns;
   *job = NULL;
   if (fl3_interpolate(P->spec->genomes, P->spec->ngenomes, P->spec->time + P->temporal_deltas[tidx], 0, cp))
      return 1;
   if (fl3_prepare_precalc_flags(cp)) {
      fprintf(stderr, "prepare xform pointers returned error: aborting.\n");
      return 1;
   }
   for (i = 0; i < cp->num_xforms; i++) fl3_xform_precalc(&cp->xform[i]);
   /* the selection table's breakpoints (what the job holds) */
   ns = cp->num_xforms - (cp->final_xform_index >= 0);
   starts = (int *)malloc(sizeof(int) * (size_t)(ns + 1) * (size_t)(ns > 0 ? ns : 1));
   if (!starts || fl3_xform_starts(cp, starts, &nrows)) {
      free(starts);
      fprintf(stderr, "create xform distrib returned error: aborting.\n");
      return 1;
   }
   /* palette scaled to the white level and the temporal weight (rect.c:778-782) */
   for (j = 0; j < 256; j++)
      for (i = 0; i < 4; i++) P->cam->dmap[j][i] = cp->palette[j].color[i] * WHITE_LEVEL * color_scalar;
   P->cam->cmap_size = P->cmap_size;
   P->cam->palette_mode = cp->palette_mode;
   if (setup_camera(P->spec, cp, P->field, P->oversample, &P->geo, P->cam)) {
      free(starts);
      return 1;
   }
   *batch_size = P->geo.sample_density * P->geo.image_width * P->geo.image_height / (P->nbatches * P->nts);
   if (first) {
      P->stats->num_iters += *batch_size;
      P->vibrancy += cp->vibrancy;
      P->gamma += cp->gamma;
      P->background[0] += cp->background[0];
      P->background[1] += cp->background[1];
      P->background[2] += cp->background[2];
      P->vib_gam_n++;
   }
   *job = fl3_pack_job_starts(cp, starts, P->cam, job_bytes);
   free(starts);
   return *job ? 0 : 1;
}

/* Production walker keys of temporal sample ts: the frame stream advances
 * by 16 words per reference thread, as rect.c:858-865 consumes it; the first
 * four seed the key. */
static void plan_keys(render_plan *P, int ts, int tidx) {
   uint64_t w[4] = {0, 0, 0, 0};
   int i, j;
   for (i = 0; i < P->nthreads; i++)
      for (j = 0; j < FLAM3_RANDSIZ; j++) {
         ub4 v = fl3_irand(&P->spec->rc);
         if (i == 0 && j < 4) w[j] = (uint64_t)(v & 0xffffffffUL);
      }
   P->keys[2 * ts] = w[0] | (w[1] << 32);
   P->keys[2 * ts + 1] = (w[2] | (w[3] << 32)) ^ ((uint64_t)tidx * 0x9E3779B97F4A7C15ULL);
}

/* Plotted samples of one production temporal sample. */
static int64_t plan_nsamples(const render_plan *P, double batch_size) {
   return (int64_t)P->nthreads * (int64_t)(batch_size / (double)P->nthreads);
}

/* Density normalisation of a batch (rect.c:933-937), from the genome of its
 * last temporal sample. */
static void plan_k1k2(const render_plan *P, double *k1, double *k2) {
   const flam3_genome *cp = &P->cp;
   const double area = P->geo.image_width * P->geo.image_height / (P->geo.ppux * P->geo.ppuy);
   *k1 = (cp->contrast * cp->brightness * PREFILTER_WHITE * 268.0 * (1.0 / (double)P->nbatches)) / 256;
   *k2 = (P->oversample * P->oversample * P->nbatches) /
         (cp->contrast * area * WHITE_LEVEL * P->geo.sample_density * P->sumfilt);
}

static void plan_de_params(const render_plan *P, fl3d_de_params *dp) {
   plan_k1k2(P, &dp->k1, &dp->k2);
   dp->oversample = P->oversample;
   dp->curve = P->cp.estimator_curve;
   dp->exact = P->replay;
}

/* Tone-map parameters (rect.c:1067-1081). */
static void plan_tm(render_plan *P, fl3d_tm_params *tp) {
   int i;
   memset(tp, 0, sizeof(*tp));
   P->vibrancy /= P->vib_gam_n;
   P->background[0] /= P->vib_gam_n / 256.0;
   P->background[1] /= P->vib_gam_n / 256.0;
   P->background[2] /= P->vib_gam_n / 256.0;
   tp->image_width = P->geo.image_width;
   tp->image_height = P->geo.image_height;
   tp->oversample = P->oversample;
   tp->de_offset = P->geo.de_offset;
   tp->filter_width = P->filter_width;
   tp->nchan = P->nchan;
   tp->transp = P->transp;
   tp->bytes_per_channel = P->bpc;
   tp->earlyclip = P->spec->earlyclip;
   tp->g = 1.0 / (P->gamma / P->vib_gam_n);
   tp->linrange = P->cp.gam_lin_thresh;
   tp->vibrancy = P->vibrancy;
   tp->highpow = P->cp.highlight_power;
   for (i = 0; i < 3; i++) tp->background[i] = P->background[i];
}

static size_t plan_out_stride(const render_plan *P) { return (size_t)P->nchan * P->bpc * P->geo.out_width; }

/* Caller-visible end of a render: stats and the optional palette strip
 * along the top (rect.c:1267-1279). */
static void plan_finish(render_plan *P) {
   int i, j;
   P->stats->badvals = P->badvals;
   if (getenv("insert_palette")) {
      int ph = 100;
      if (ph >= P->geo.image_height) ph = P->geo.image_height;
      for (j = 0; j < ph; j++)
         for (i = 0; i < P->geo.image_width; i++) {
            unsigned char *p = (unsigned char *)P->out + P->nchan * (i + j * P->geo.out_width);
            p[0] = (unsigned char)P->cam->dmap[i * 256 / P->geo.image_width][0];
            p[1] = (unsigned char)P->cam->dmap[i * 256 / P->geo.image_width][1];
            p[2] = (unsigned char)P->cam->dmap[i * 256 / P->geo.image_width][2];
         }
   }
   P->stats->render_seconds = (int)(time(NULL) - P->tstart);
}

/* Jobs of temporal samples [t0, t0 + n) built on several host threads,
 * each with its own genome and camera scratch: what plan_ts_job(P, t0 + k,
 * 0, ...) returns for each, and the genome values the first pass adds to
 * the tone-map averages.  P itself is not changed. */
typedef struct {
   void *job;
   size_t bytes;
   double batch_size, vibrancy, gamma, background[3];
   int ok;
} ts_result;

typedef struct {
   const render_plan *P;
   int t0, n, k0, step;
   ts_result *res;
} ts_arg;

static void *ts_thread(void *p) {
   ts_arg *a = (ts_arg *)p;
   render_plan T = *a->P;   /* shallow: shared read-only tables, private scratch below */
   stat_struct st;
   int k;
   memset(&T.cp, 0, sizeof(T.cp));
   T.cam = (fl3h_camera *)calloc(1, sizeof(fl3h_camera));
   memset(&st, 0, sizeof(st));
   T.stats = &st;
   for (k = a->k0; k < a->n; k += a->step) {
      ts_result *r = &a->res[k];
      r->ok = T.cam && plan_ts_job(&T, a->t0 + k, 0, &r->job, &r->bytes, &r->batch_size) == 0;
      if (!r->ok) continue;
      r->vibrancy = T.cp.vibrancy;
      r->gamma = T.cp.gamma;
      r->background[0] = T.cp.background[0];
      r->background[1] = T.cp.background[1];
      r->background[2] = T.cp.background[2];
   }
   free(T.cam);
   fl3_genome_free(&T.cp);
   return NULL;
}

static int build_ts_jobs(render_plan *P, int t0, int n, ts_result *res) {
   pthread_t th[FL3_PREP_THREADS];
   ts_arg ta[FL3_PREP_THREADS];
   int nt = n < FL3_PREP_THREADS ? n : FL3_PREP_THREADS, i, started = 0;
   for (i = 0; i < nt; i++) {
      ta[i].P = P;
      ta[i].t0 = t0;
      ta[i].n = n;
      ta[i].k0 = i;
      ta[i].step = nt;
      ta[i].res = res;
   }
   for (i = 1; i < nt; i++)
      if (pthread_create(&th[i], NULL, ts_thread, &ta[i]) == 0)
         started |= 1 << i;
      else
         ts_thread(&ta[i]);
   if (nt > 0) ts_thread(&ta[0]);
   for (i = 1; i < nt; i++)
      if (started & (1 << i)) pthread_join(th[i], NULL);
   for (i = 0; i < n; i++)
      if (!res[i].ok) return 1;
   return 0;
}

int fl3h_render(flam3_frame *spec, void *out, int field, int nchan, int transp, stat_struct *stats,
                const fl3h_render_opts *opts) {
   render_plan P;
   int batch_num, ts, i, rv = 1, aborted = 0, nts;
   void **bjobs = NULL;   /* temporal samples waiting for their launch */
   size_t *bbytes = NULL;
   int64_t *bns = NULL;   /* and their sample counts */
   int bn = 0, bts0 = 0;

   if (plan_open(&P, spec, out, field, nchan, transp, stats, opts)) return 1;
   nts = P.nts;
   if (fl3d_begin(P.geo.width, P.geo.height, spec->bits)) {
      fprintf(stderr, "flam3_amd: %s\n", fl3d_error());
      goto fail;
   }
   bjobs = (void **)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(void *));
   bbytes = (size_t *)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(size_t));
   bns = (int64_t *)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(int64_t));
   if (!bjobs || !bbytes || !bns) goto fail;

   for (batch_num = 0; batch_num < P.nbatches && !aborted; batch_num++) {
      fl3_de_filters de;
      double batch_bad = 0.0;
      int saturated = 0, exact = 0, pass;

      if (plan_batch_de(&P, batch_num, &de)) goto fail;
      /* every batch starts from empty buckets; its density estimate adds to
       * the accumulator (rect.c:709-721) */
      if (batch_num > 0 && fl3d_clear_buckets()) goto devfail_de;

      /* Pass 0 uses fast deposits; if a hit counter saturated, pass 1
       * repeats the batch with saturating adds on every channel, reusing
       * the walker keys so the same points are drawn. */
      for (pass = 0; pass < 2; pass++) {
         if (pass == 1) {
            if (!saturated || P.replay) break;
            exact = 1;
            P.badvals -= batch_bad;
            batch_bad = 0.0;
            if (fl3d_clear_buckets()) goto devfail_de;
         }
         ts = 0;
         if (pass == 0 && !P.replay && !spec->progress && nts >= 4) {
            /* temporal samples built on several threads, then launched in
             * order exactly as below */
            ts_result *res = (ts_result *)calloc((size_t)nts, sizeof(ts_result));
            int k, failed;
            void *last = NULL;
            size_t lb;
            double lbs;
            if (!res) {
               fl3_free_de_filters(&de);
               goto fail;
            }
            failed = build_ts_jobs(&P, batch_num * nts, nts, res);
            /* the plan's genome, camera and geometry as the last sample left them */
            if (!failed) failed = plan_ts_job(&P, batch_num * nts + nts - 1, 0, &last, &lb, &lbs);
            free(last);
            for (ts = 0; !failed && ts < nts; ts++) {
               const ts_result *r = &res[ts];
               double bad = 0.0;
               int sat = 0;
               int64_t nsamp;
               stats->num_iters += r->batch_size;
               P.vibrancy += r->vibrancy;
               P.gamma += r->gamma;
               P.background[0] += r->background[0];
               P.background[1] += r->background[1];
               P.background[2] += r->background[2];
               P.vib_gam_n++;
               plan_keys(&P, ts, batch_num * nts + ts);
               nsamp = plan_nsamples(&P, r->batch_size);
               if (bn == max_batch_jobs(opts)) {
                  if (flush_batch(bjobs, bbytes, &bn, bns, P.keys + 2 * bts0, exact, opts, &bad, &sat)) {
                     failed = 2;
                     break;
                  }
                  if (sat) saturated = 1;
               }
               if (bn == 0) bts0 = ts;
               bjobs[bn] = r->job;
               bbytes[bn] = r->bytes;
               bns[bn] = nsamp;
               bn++;
               res[ts].job = NULL;
               P.badvals += bad;
               batch_bad += bad;
            }
            for (k = 0; k < nts; k++) free(res[k].job);
            free(res);
            if (failed == 2) goto devfail_de;
            if (failed) {
               fl3_free_de_filters(&de);
               goto fail;
            }
            ts = nts;
         }
         for (; ts < nts; ts++) {
            const int tidx = batch_num * nts + ts;
            void *job;
            size_t job_bytes;
            double batch_size, bad = 0.0;
            int sat = 0;

            if (plan_ts_job(&P, tidx, pass == 0, &job, &job_bytes, &batch_size)) {
               fl3_free_de_filters(&de);
               goto fail;
            }
            if (poll_progress(spec, 100.0 * (((double)ts / nts) + batch_num) / P.nbatches, 0, 0.0)) {
               free(job);
               aborted = 1;
               break;
            }

            if (P.replay) {
               /* walker w = reference thread w (rect.c:858-865) */
               for (i = 0; i < P.nthreads; i++) thread_state(&spec->rc, P.walker_states + 36 * i);
               if (fl3d_replay(job, job_bytes, P.walker_states, P.nthreads, batch_size / (double)P.nthreads,
                               spec->sub_batch_size > 0 ? spec->sub_batch_size : 10000,
                               spec->earlyclip ? FUSE_EARLYCLIP : FUSE_NORMAL, &bad)) {
                  free(job);
                  goto devfail_de;
               }
            } else {
               int64_t nsamp;
               if (pass == 0) plan_keys(&P, ts, tidx);
               /* Temporal samples go to the GPU together, one launch for up
                * to max_batch_jobs of them. */
               nsamp = plan_nsamples(&P, batch_size);
               if (bn == max_batch_jobs(opts)) {
                  if (flush_batch(bjobs, bbytes, &bn, bns, P.keys + 2 * bts0, exact, opts, &bad, &sat))
                     goto devfail_de;
                  if (sat) saturated = 1;
               }
               if (bn == 0) bts0 = ts;
               bjobs[bn] = job;
               bbytes[bn] = job_bytes;
               bns[bn] = nsamp;
               bn++;
               job = NULL;
            }
            free(job);
            P.badvals += bad;
            batch_bad += bad;
         }
         if (!aborted && bn > 0) {
            double bad = 0.0;
            int sat = 0;
            if (flush_batch(bjobs, bbytes, &bn, bns, P.keys + 2 * bts0, exact, opts, &bad, &sat)) goto devfail_de;
            P.badvals += bad;
            batch_bad += bad;
            if (sat) saturated = 1;
         }
         if (aborted) break;
      }
      if (!aborted) {
         if (de.max_filter_index == 0) {
            double k1, k2;
            plan_k1k2(&P, &k1, &k2);
            if (fl3d_logscale(k1, k2)) goto devfail_de;
         } else {
            fl3d_de_params dp;
            plan_de_params(&P, &dp);
            if (fl3d_density(&dp, &de)) goto devfail_de;
            if (poll_progress(spec, 100.0, 1, 0.0)) aborted = 1;
         }
      }
      fl3_free_de_filters(&de);
      continue;
   devfail_de:
      fl3_free_de_filters(&de);
      goto devfail;
   }

   if (!aborted) {
      fl3d_tm_params tp;
      plan_tm(&P, &tp);
      if (fl3d_tonemap(&tp, P.filter, P.out, plan_out_stride(&P))) goto devfail;
      plan_finish(&P);
   } else {
      stats->badvals = P.badvals;
      stats->render_seconds = (int)(time(NULL) - P.tstart);
   }
   rv = 0;
   goto cleanup;

devfail:
   fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
fail:
   rv = 1;
cleanup:
   for (i = 0; i < bn; i++) free(bjobs[i]);
   free(bjobs);
   free(bbytes);
   free(bns);
   plan_close(&P);
   return rv;
}

/* ---------------------------------------------------------------- many renders
 *
 * Several renders in two phases: fl3_many_prepare does the host work of
 * every render (plans, temporal-sample jobs, walker keys, DE filters) and
 * fl3_many_execute the device work, so a caller can prepare the next set
 * while the GPU runs this one.  Small renders share the GPU: the walkers of
 * every temporal sample of every small render run in one launch, each
 * depositing into its own render's histogram, then each render's density
 * estimation and tone map run and all their images come back in one copy.
 * Large renders run one after another with fl3h_render's launches.  Every
 * render draws the walkers, keys and samples fl3h_render would draw for it
 * alone, and its integer deposits do not depend on order, so its image is
 * the one fl3h_render makes. */

/* Production renders of one batch without a progress callback are
 * prepared here.  1 = shared (its walkers join the shared launch, with the
 * bits of the other shared renders); 2 = alone, with fl3h_render's
 * launches.  0: fl3h_render itself. */
static int plan_kind(const render_plan *P, const fl3h_render_opts *opts, int bits) {
   double est;
   if (P->nbatches != 1 || P->replay || P->spec->progress) return 0;
   if (opts && opts->deposit == 1) return 0;  /* forced atomics may saturate: fl3h_render's second pass */
   if (opts && opts->deposit == 2) return 2;
   if (bits && P->spec->bits != bits) return 2;
   /* a stepped palette at bits 32/33: the shared launch bins job records,
    * which sums the same integers as fl3h_render's deposit of any size (a
    * linear palette's records would round the blend fraction) */
   if (P->cp.palette_mode == 0 && P->spec->bits != 64 &&
       (long long)P->geo.width * P->geo.height <= FL3_SHARED_MAX_CELLS)
      return 1;
   /* else small renders, whose shared deposits are per-sample atomics as
    * fl3h_render's are */
   est = P->cp.sample_density * pow(2.0, 2.0 * P->cp.zoom) * P->geo.image_width * P->geo.image_height;
   if (est < (double)FL3_SHARED_MAX_SAMPLES) return 1;
   return 2;
}

/* No cell of the render can reach 2^32 in any channel, so the fast deposit
 * never saturates (fl3h_render's saturating second pass is never needed);
 * otherwise its jobs deposit with saturating adds. */
static int jobs_cannot_saturate(const fl3d_multi_job *m, int n) {
   double total = 0.0, maxd = 255.0;
   int j, e;
   for (j = 0; j < n; j++) {
      const fl3_job_hdr *h = (const fl3_job_hdr *)m[j].job;
      const double *dm = (const double *)((const unsigned char *)h + h->off_dmap);
      for (e = 0; e < 1024; e++) maxd = dm[e] > maxd ? dm[e] : maxd;
      total += (double)m[j].nsamples;
   }
   return total * (maxd + 1.0) < 4294967295.0;
}

struct fl3_many {
   fl3h_request *reqs;
   int n;
   fl3h_render_opts opts;
   int has_opts;
   render_plan *P;
   fl3_de_filters *de;
   int *kind;              /* per request: -1 failed, 0 fl3h_render, 1 small, 2 large */
   int *job0, *njob;       /* per request: its jobs in mj */
   int *live;              /* small renders in launch order */
   int *one_pal;           /* per request: its jobs share 4-byte records */
   int nlive, bits;
   fl3d_multi_job *mj;
   int nmj;
   fl3d_multi_job **rj;    /* per request, before they go into mj */
   fl3_delivery dv;        /* hand_over NULL: copy the images in place */
};

struct fl3_images {
   fl3_many *B;
   const unsigned char *host;   /* the downloaded arena; NULL if the download failed */
   int slot, n;
   fl3d_image_dst *dst;         /* n destinations */
   int *req;                    /* n request indices */
};

void fl3_many_set_delivery(fl3_many *B, const fl3_delivery *d) {
   if (B && d) B->dv = *d;
}

/* Rows of each image from the downloaded arena into the caller's buffer. */
static void scatter_images(const unsigned char *host, const fl3d_image_dst *d, int n) {
   int i, j;
   for (i = 0; i < n; i++) {
      const unsigned char *src = host + d[i].img_off;
      unsigned char *dst = (unsigned char *)d[i].out;
      if (d[i].out_stride == d[i].row_bytes)
         memcpy(dst, src, d[i].row_bytes * (size_t)d[i].rows);
      else
         for (j = 0; j < d[i].rows; j++)
            memcpy(dst + (size_t)j * d[i].out_stride, src + (size_t)j * d[i].row_bytes, d[i].row_bytes);
   }
}

int fl3_images_deliver(fl3_images *im) {
   int slot, i;
   if (!im) return -1;
   if (im->host) {
      scatter_images(im->host, im->dst, im->n);
      for (i = 0; i < im->n; i++) {
         plan_finish(&im->B->P[im->req[i]]);
         im->B->reqs[im->req[i]].rc = 0;
      }
   }
   slot = im->slot;
   free(im->dst);
   free(im->req);
   free(im);
   return slot;
}

void fl3_many_free(fl3_many *B) {
   int i;
   if (!B) return;
   for (i = 0; i < B->nmj; i++) free((void *)B->mj[i].job);
   if (B->rj)
      for (i = 0; i < B->n; i++)
         if (B->rj[i]) {
            int k;
            for (k = 0; k < B->njob[i]; k++) free((void *)B->rj[i][k].job);
            free(B->rj[i]);
         }
   free(B->rj);
   if (B->P)
      for (i = 0; i < B->n; i++) {
         if (B->de) fl3_free_de_filters(&B->de[i]);
         plan_close(&B->P[i]);
      }
   free(B->mj);
   free(B->one_pal);
   free(B->live);
   free(B->njob);
   free(B->job0);
   free(B->kind);
   free(B->de);
   free(B->P);
   free(B);
}

/* Walkers per run of up to max_batch_jobs consecutive temporal samples, as
 * fl3h_render launches them (sized by the run's largest sample count). */
static int assign_walkers(fl3d_multi_job *m, int n, const fl3h_render_opts *opts) {
   int ts = 0, i;
   while (ts < n) {
      int e = ts;
      long long w;
      int64_t nmax = 0;
      while (e < n && e - ts < max_batch_jobs(opts)) {
         if (m[e].nsamples > nmax) nmax = m[e].nsamples;
         e++;
      }
      w = fl3d_job_walkers(nmax, e - ts, opts ? opts->nwalkers : 0);
      if (w < 256) return 1;
      for (i = ts; i < e; i++) {
         m[i].walkers = (int)w;
         m[i].run = ts;
      }
      ts = e;
   }
   return 0;
}

/* Host work of request r: plan, DE filters, temporal-sample jobs with their
 * keys (the request's frame stream advances here) and walker counts.
 * Leaves kind[r] -1 (failed), 0 (fl3h_render later) or 1 / 2 with the jobs
 * in rj[r]. */
static void prep_one(fl3_many *B, int r, const fl3h_render_opts *opts) {
   fl3h_request *q = &B->reqs[r];
   render_plan *Q = &B->P[r];
   fl3d_multi_job *m;
   randctx rc0;
   int ts, i;
   q->rc = 1;
   B->kind[r] = -1;
   if (plan_open(Q, q->frame, q->out, q->field, q->nchan, q->transp, q->stats, opts)) return;
   B->kind[r] = plan_kind(Q, opts, 0);
   if (B->kind[r] == 0) {
      plan_close(Q);
      return;
   }
   rc0 = q->frame->rc;
   m = (fl3d_multi_job *)calloc((size_t)Q->nts, sizeof(fl3d_multi_job));
   if (!m || plan_batch_de(Q, 0, &B->de[r])) {
      free(m);
      B->kind[r] = -1;
      plan_close(Q);
      return;
   }
   for (ts = 0; ts < Q->nts; ts++) {
      double batch_size;
      void *job;
      plan_keys(Q, ts, ts);
      if (plan_ts_job(Q, ts, 1, &job, &m[ts].bytes, &batch_size)) break;
      m[ts].job = job;
      m[ts].nsamples = plan_nsamples(Q, batch_size);
      m[ts].key[0] = Q->keys[2 * ts];
      m[ts].key[1] = Q->keys[2 * ts + 1];
   }
   if (ts == Q->nts && !jobs_cannot_saturate(m, ts))
      for (i = 0; i < ts; i++) m[i].exact = 1;
   if (ts == Q->nts && assign_walkers(m, ts, opts)) {
      fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
      ts = -1;
   }
   if (ts != Q->nts) {
      for (i = 0; i < Q->nts; i++) free((void *)m[i].job);
      free(m);
      q->frame->rc = rc0;
      B->kind[r] = -1;
      fl3_free_de_filters(&B->de[r]);
      plan_close(Q);
      return;
   }
   B->rj[r] = m;
   B->njob[r] = Q->nts;
   B->one_pal[r] = 1;
   for (ts = 0; ts < Q->nts && B->one_pal[r]; ts++) B->one_pal[r] = fl3d_jobs_share_palette(m[0].job, m[ts].job);
}

typedef struct {
   fl3_many *B;
   const fl3h_render_opts *opts;
   int r0, step;
} prep_arg;

static void *prep_thread(void *p) {
   prep_arg *a = (prep_arg *)p;
   int r;
   for (r = a->r0; r < a->B->n; r += a->step) prep_one(a->B, r, a->opts);
   return NULL;
}

fl3_many *fl3_many_prepare(fl3h_request *reqs, int n, const fl3h_render_opts *opts) {
   fl3_many *B = (fl3_many *)calloc(1, sizeof(fl3_many));
   int r, i, nt;
   if (!B) return NULL;
   B->reqs = reqs;
   B->n = n;
   if (opts) {
      B->opts = *opts;
      B->has_opts = 1;
   }
   opts = B->has_opts ? &B->opts : NULL;
   B->P = (render_plan *)calloc((size_t)(n > 0 ? n : 1), sizeof(render_plan));
   B->de = (fl3_de_filters *)calloc((size_t)(n > 0 ? n : 1), sizeof(fl3_de_filters));
   B->kind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
   B->job0 = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
   B->njob = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
   B->live = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
   B->rj = (fl3d_multi_job **)calloc((size_t)(n > 0 ? n : 1), sizeof(fl3d_multi_job *));
   B->one_pal = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
   if (!B->P || !B->de || !B->kind || !B->job0 || !B->njob || !B->live || !B->rj || !B->one_pal) {
      fl3_many_free(B);
      return NULL;
   }
   nt = n < FL3_PREP_THREADS ? n : FL3_PREP_THREADS;
   if (nt > 1) {
      pthread_t th[FL3_PREP_THREADS];
      prep_arg pa[FL3_PREP_THREADS];
      int started = 0;
      for (i = 0; i < nt; i++) {
         pa[i].B = B;
         pa[i].opts = opts;
         pa[i].r0 = i;
         pa[i].step = nt;
      }
      /* thread 0's share runs here; a thread that fails to start leaves
       * its share to this thread too */
      for (i = 1; i < nt; i++)
         if (pthread_create(&th[i], NULL, prep_thread, &pa[i]) == 0)
            started |= 1 << i;
         else
            prep_thread(&pa[i]);
      prep_thread(&pa[0]);
      for (i = 1; i < nt; i++)
         if (started & (1 << i)) pthread_join(th[i], NULL);
   } else
      for (r = 0; r < n; r++) prep_one(B, r, opts);
   /* in request order: the shared renders take the bits of the first of
    * them (others run alone), and the jobs go into one array */
   for (r = 0; r < n; r++) {
      const int nj = B->njob[r];
      fl3d_multi_job *t;
      if (B->kind[r] < 1) continue;
      if (B->kind[r] == 1) {
         if (!B->bits)
            B->bits = B->P[r].spec->bits;
         else if (B->P[r].spec->bits != B->bits)
            B->kind[r] = 2;
      }
      t = (fl3d_multi_job *)realloc(B->mj, sizeof(fl3d_multi_job) * (size_t)(B->nmj + nj));
      if (!t) {
         B->kind[r] = -1;
         continue;
      }
      B->mj = t;
      memcpy(B->mj + B->nmj, B->rj[r], sizeof(fl3d_multi_job) * (size_t)nj);
      free(B->rj[r]);
      B->rj[r] = NULL;
      B->job0[r] = B->nmj;
      B->nmj += nj;
      if (B->kind[r] == 1) {
         for (i = 0; i < nj; i++) B->mj[B->job0[r] + i].render = B->nlive;
         B->live[B->nlive++] = r;
      }
   }
   return B;
}

/* Density estimation (or log scaling) of the selected render. */
static int many_density(render_plan *Q, fl3_de_filters *de) {
   if (de->max_filter_index == 0) {
      double k1, k2;
      plan_k1k2(Q, &k1, &k2);
      return fl3d_logscale(k1, k2);
   } else {
      fl3d_de_params dp;
      plan_de_params(Q, &dp);
      return fl3d_density(&dp, de);
   }
}

/* Shared renders live[i0, i1): one histogram allocation, one walker launch
 * (per kernel tier), their density estimation and tone maps, one image
 * download. */
static int many_group(fl3_many *B, int i0, int i1) {
   int *widths, *heights, *sat, i, ok = 0;
   double *bad;
   fl3d_multi_job *mj;
   fl3d_image_dst *dst;
   size_t img_total = 0, filt_total = 0;
   const int nl = i1 - i0;
   int nm = 0;
   if (nl <= 0) return 0;
   widths = (int *)malloc(sizeof(int) * (size_t)nl);
   heights = (int *)malloc(sizeof(int) * (size_t)nl);
   bad = (double *)calloc((size_t)nl, sizeof(double));
   sat = (int *)calloc((size_t)nl, sizeof(int));
   dst = (fl3d_image_dst *)calloc((size_t)nl, sizeof(fl3d_image_dst));
   mj = (fl3d_multi_job *)malloc(sizeof(fl3d_multi_job) * (size_t)(B->nmj > 0 ? B->nmj : 1));
   if (widths && heights && bad && sat && dst && mj) {
      for (i = 0; i < nl; i++) {
         const int r = B->live[i0 + i];
         const render_plan *Q = &B->P[r];
         int k;
         widths[i] = Q->geo.width;
         heights[i] = Q->geo.height;
         dst[i].img_off = img_total;
         dst[i].out = Q->out;
         dst[i].out_stride = plan_out_stride(Q);
         dst[i].row_bytes = (size_t)Q->geo.image_width * Q->nchan * Q->bpc;
         dst[i].rows = Q->geo.image_width > 0 ? Q->geo.image_height : 0;
         img_total += (dst[i].row_bytes * (size_t)Q->geo.image_height + 255) & ~(size_t)255;
         filt_total += (sizeof(double) * Q->filter_width * Q->filter_width + 255) & ~(size_t)255;
         memcpy(mj + nm, B->mj + B->job0[r], sizeof(fl3d_multi_job) * (size_t)B->njob[r]);
         for (k = 0; k < B->njob[r]; k++) mj[nm + k].render = i;
         nm += B->njob[r];
      }
      ok = fl3d_begin_many(widths, heights, nl, B->bits) == 0 &&
           fl3d_iterate_multi(mj, nm, FUSE_WALKER, 0, nl, bad, sat) == 0;
      if (ok) {
         /* density estimation (or log scaling) of the group in one launch pair */
         fl3d_de_params *dps = (fl3d_de_params *)calloc((size_t)nl, sizeof(fl3d_de_params));
         const fl3_de_filters **des = (const fl3_de_filters **)calloc((size_t)nl, sizeof(fl3_de_filters *));
         ok = dps && des;
         for (i = 0; ok && i < nl; i++) {
            render_plan *Q = &B->P[B->live[i0 + i]];
            const fl3_de_filters *d = &B->de[B->live[i0 + i]];
            Q->badvals = bad[i];
            plan_de_params(Q, &dps[i]);
            des[i] = d->max_filter_index > 0 ? d : NULL;
         }
         if (ok) ok = fl3d_density_many(dps, des, nl) == 0;
         free(dps);
         free((void *)des);
      }
      if (ok) ok = fl3d_images_begin(img_total, filt_total) == 0;
      if (ok) {
         /* the group's tone maps in one launch, the images in one copy */
         fl3d_tm_params *tps = (fl3d_tm_params *)calloc((size_t)nl, sizeof(fl3d_tm_params));
         const double **fls = (const double **)calloc((size_t)nl, sizeof(double *));
         size_t *offs = (size_t *)calloc((size_t)nl, sizeof(size_t));
         ok = tps && fls && offs;
         for (i = 0; ok && i < nl; i++) {
            render_plan *Q = &B->P[B->live[i0 + i]];
            plan_tm(Q, &tps[i]);
            fls[i] = Q->filter;
            offs[i] = dst[i].img_off;
         }
         if (ok) ok = fl3d_tonemap_many(tps, fls, offs, nl) == 0;
         free(offs);
         free((void *)fls);
         free(tps);
      }
      if (ok && B->dv.hand_over) {
         /* download into a free slot; the copies into the callers' buffers
          * happen on the delivering thread */
         fl3_images *im = (fl3_images *)calloc(1, sizeof(fl3_images));
         int *req = (int *)malloc(sizeof(int) * (size_t)nl);
         ok = im && req;
         if (ok) {
            im->B = B;
            im->slot = B->dv.acquire(B->dv.ctx);
            im->n = nl;
            im->dst = dst;
            im->req = req;
            for (i = 0; i < nl; i++) req[i] = B->live[i0 + i];
            dst = NULL;
            req = NULL;
            if (fl3d_images_download(im->slot, img_total, &im->host)) {
               ok = 0;
               im->host = NULL;
            }
            B->dv.hand_over(B->dv.ctx, im);   /* also a failed one: it frees the slot */
            im = NULL;
         }
         free(im);
         free(req);
         if (!ok) fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
      } else {
         if (ok) ok = fl3d_images_fetch(dst, nl) == 0;
         if (!ok) fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
         for (i = 0; ok && i < nl; i++) {
            plan_finish(&B->P[B->live[i0 + i]]);
            B->reqs[B->live[i0 + i]].rc = 0;
         }
      }
   }
   free(mj);
   free(dst);
   free(sat);
   free(bad);
   free(heights);
   free(widths);
   return ok ? 0 : 1;
}

/* The shared renders in groups whose joint cell space fits the tile-binned
 * deposit (at most FL3_SHARED_MAX_CELLS cells, 65535 jobs); renders that all
 * share one palette stay below 2^24 cells together, so that their records
 * take 4 bytes. */
static int many_small(fl3_many *B) {
   int i0 = 0, rc = 0;
   while (i0 < B->nlive) {
      long long cells = 0;
      int jobs = 0, i1 = i0, one = 1;
      const void *job0 = B->mj[B->job0[B->live[i0]]].job;
      while (i1 < B->nlive) {
         const int r = B->live[i1];
         const render_plan *Q = &B->P[r];
         const long long c = (long long)Q->geo.width * Q->geo.height;
         const int nj = B->njob[r];
         const int o = one && B->one_pal[r] && fl3d_jobs_share_palette(job0, B->mj[B->job0[r]].job);
         const long long cap = o ? FL3_SHARED_NARROW_CELLS : FL3_SHARED_MAX_CELLS;
         if (i1 > i0 && (cells + c > cap || jobs + nj > 65535)) break;
         one = o;
         cells += c;
         jobs += nj;
         i1++;
      }
      rc |= many_group(B, i0, i1);
      i0 = i1;
   }
   return rc;
}

/* A large prepared render: fl3h_render's launches (runs of up to
 * max_batch_jobs temporal samples, each with its own sample count), density
 * estimation and tone map. */
static int many_large(fl3_many *B, int r) {
   render_plan *Q = &B->P[r];
   const fl3h_render_opts *opts = B->has_opts ? &B->opts : NULL;
   fl3d_multi_job *m = B->mj + B->job0[r];
   int exact, ts, e, i, saturated = 0;
   fl3d_tm_params tp;
   if (fl3d_begin(Q->geo.width, Q->geo.height, Q->spec->bits)) goto devfail;
   for (exact = 0; exact < 2; exact++) {
      if (exact) {
         if (!saturated) break;
         if (fl3d_clear_buckets()) goto devfail;
      }
      Q->badvals = 0.0;
      for (ts = 0; ts < B->njob[r]; ts = e) {
         const void *jobs[MAX_BATCH_JOBS];
         size_t bytes[MAX_BATCH_JOBS];
         uint64_t keys[2 * MAX_BATCH_JOBS];
         int64_t ns[MAX_BATCH_JOBS], nmax = 0;
         int same = 1;
         fl3d_iter_params ip;
         double bad = 0.0;
         int sat = 0;
         for (e = ts; e < B->njob[r] && e - ts < max_batch_jobs(opts); e++) {
            jobs[e - ts] = m[e].job;
            bytes[e - ts] = m[e].bytes;
            keys[2 * (e - ts)] = m[e].key[0];
            keys[2 * (e - ts) + 1] = m[e].key[1];
            ns[e - ts] = m[e].nsamples;
            if (m[e].nsamples > nmax) nmax = m[e].nsamples;
            if (m[e].nsamples != m[ts].nsamples) same = 0;
         }
         memset(&ip, 0, sizeof(ip));
         ip.fuse = FUSE_WALKER;
         ip.nsamples = nmax;
         ip.job_nsamples = same ? NULL : ns;
         ip.key[0] = keys[0];
         ip.key[1] = keys[1];
         ip.exact = exact;
         ip.nwalkers = opts ? opts->nwalkers : 0;
         ip.deposit = opts ? opts->deposit : 0;
         ip.chunk = opts ? opts->chunk_samples : 0;
         if (fl3d_iterate_batch(jobs, bytes, e - ts, &ip, keys, &bad, &sat)) goto devfail;
         Q->badvals += bad;
         if (sat) saturated = 1;
      }
   }
   if (many_density(Q, &B->de[r])) goto devfail;
   plan_tm(Q, &tp);
   if (fl3d_tonemap(&tp, Q->filter, Q->out, plan_out_stride(Q))) goto devfail;
   plan_finish(Q);
   B->reqs[r].rc = 0;
   (void)i;
   return 0;
devfail:
   fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
   return 1;
}

int fl3_many_execute(fl3_many *B) {
   int r, failures = 0;
   const fl3h_render_opts *opts;
   if (!B) return 0;
   opts = B->has_opts ? &B->opts : NULL;
   many_small(B);
   for (r = 0; r < B->n; r++) {
      fl3h_request *q = &B->reqs[r];
      if (B->kind[r] == 2)
         many_large(B, r);
      else if (B->kind[r] == 0)
         q->rc = fl3h_render(q->frame, q->out, q->field, q->nchan, q->transp, q->stats, opts);
   }
   for (r = 0; r < B->n; r++) failures += B->reqs[r].rc != 0;
   return failures;
}

int fl3h_render_many(fl3h_request *reqs, int n, const fl3h_render_opts *opts) {
   fl3_many *B;
   int failures, r;
   if (n <= 0) return 0;
   B = fl3_many_prepare(reqs, n, opts);
   if (!B) {
      for (r = 0; r < n; r++) reqs[r].rc = 1;
      return n;
   }
   failures = fl3_many_execute(B);
   fl3_many_free(B);
   return failures;
}

int flam3_render(flam3_frame *f, void *out, int field, int nchan, int transp, stat_struct *stats) {
   return fl3h_render(f, out, field, nchan, transp, stats, NULL);
}

/* ---------------------------------------------------------------- helpers */

/* Seed the frame's ISAAC context from a string, as flam3_init_frame does
 * with the isaac_seed environment variable (flam3.c:2494-2516); NULL seeds
 * from the time. */
void fl3h_init_frame(flam3_frame *f, const char *isaac_seed) {
   int k;
   memset(f->rc.randrsl, 0, sizeof(f->rc.randrsl));
   if (isaac_seed == NULL) {
      long t = (long)time(0);
      for (k = 0; k < FLAM3_RANDSIZ; k++) f->rc.randrsl[k] = (ub4)t;
   } else {
      strncpy((char *)&f->rc.randrsl, isaac_seed, FLAM3_RANDSIZ * sizeof(ub4));
   }
   fl3_isaac_init(&f->rc, 1);
}

/* A genome with the parser's defaults (clear_cp with defaults on,
 * flam3.c:1270) and nstd standard xforms plus an optional final xform, each
 * with every variation weight at 0 as parse_flame_element leaves them
 * (parser.c:908-935). */
flam3_genome *fl3h_genome_new(int nstd, int has_final) {
   flam3_genome *g = (flam3_genome *)calloc(1, sizeof(flam3_genome));
   int i;
   if (!g) return NULL;
   g->final_xform_index = -1;
   fl3_genome_defaults(g, 1);
   for (i = 0; i < 256; i++) g->palette[i].index = -1;
   if (nstd > 0 && fl3_genome_add_xforms(g, nstd, 0, 0)) goto bad;
   for (i = 0; i < nstd; i++) g->xform[i].var[0] = 0.0;
   if (has_final) {
      if (fl3_genome_add_xforms(g, 1, 0, 1)) goto bad;
      g->xform[g->final_xform_index].var[0] = 0.0;
   }
   return g;
bad:
   fl3_genome_free(g);
   free(g);
   return NULL;
}

void fl3h_genome_free(flam3_genome *g) {
   if (!g) return;
   fl3_genome_free(g);
   free(g);
}
