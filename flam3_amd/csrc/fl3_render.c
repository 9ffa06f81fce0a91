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

/* Job records carry the job index in 16 bits. */
#define MAX_BATCH_JOBS 4096
/* Renders below this many plotted samples may share walker launches. */
#define FL3_SHARED_MAX_SAMPLES (1LL << 24)

/* One launch over the *n temporal samples collected so far (nsamples each,
 * keys[2t], keys[2t+1] for the t-th); frees their jobs and empties the set. */
static int flush_batch(void **jobs, size_t *bytes, int *n, int64_t nsamples, const uint64_t *keys, int exact,
                       const fl3h_render_opts *opts, double *bad, int *sat) {
   fl3d_iter_params ip;
   int rc, t;
   memset(&ip, 0, sizeof(ip));
   /* The reference fuses 15 steps (100 with earlyclip) before each
    * 10000-sample sub-batch; a production walker fuses once before a few
    * hundred to a few thousand samples, so it fuses 100 steps to keep the
    * start transient out of the image. */
   ip.fuse = FUSE_WALKER;
   ip.nsamples = nsamples;
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
   unsigned short *distrib;
   int i, j, nrows;
   *job = NULL;
   if (fl3_interpolate(P->spec->genomes, P->spec->ngenomes, P->spec->time + P->temporal_deltas[tidx], 0, cp))
      return 1;
   if (fl3_prepare_precalc_flags(cp)) {
      fprintf(stderr, "prepare xform pointers returned error: aborting.\n");
      return 1;
   }
   for (i = 0; i < cp->num_xforms; i++) fl3_xform_precalc(&cp->xform[i]);
   distrib = fl3_create_xform_distrib(cp, &nrows);
   if (!distrib) {
      fprintf(stderr, "create xform distrib returned error: aborting.\n");
      return 1;
   }
   /* palette scaled to the white level and the temporal weight (rect.c:778-782) */
   for (j = 0; j < 256; j++)
      for (i = 0; i < 4; i++) P->cam->dmap[j][i] = cp->palette[j].color[i] * WHITE_LEVEL * color_scalar;
   P->cam->cmap_size = P->cmap_size;
   P->cam->palette_mode = cp->palette_mode;
   if (setup_camera(P->spec, cp, P->field, P->oversample, &P->geo, P->cam)) {
      free(distrib);
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
   *job = fl3_pack_job(cp, distrib, P->cam, job_bytes);
   free(distrib);
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
   dp->exact = P->replay || P->spec->bits == 64;
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

int fl3h_render(flam3_frame *spec, void *out, int field, int nchan, int transp, stat_struct *stats,
                const fl3h_render_opts *opts) {
   render_plan P;
   int batch_num, ts, i, rv = 1, aborted = 0, nts;
   void **bjobs = NULL;   /* temporal samples waiting for their launch */
   size_t *bbytes = NULL;
   int bn = 0, bts0 = 0;
   int64_t bns = 0;

   if (plan_open(&P, spec, out, field, nchan, transp, stats, opts)) return 1;
   nts = P.nts;
   if (fl3d_begin(P.geo.width, P.geo.height, spec->bits)) {
      fprintf(stderr, "flam3_amd: %s\n", fl3d_error());
      goto fail;
   }
   bjobs = (void **)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(void *));
   bbytes = (size_t *)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(size_t));
   if (!bjobs || !bbytes) goto fail;

   for (batch_num = 0; batch_num < P.nbatches && !aborted; batch_num++) {
      fl3_de_filters de;
      double batch_bad = 0.0;
      int saturated = 0, exact = 0, pass;

      if (plan_batch_de(&P, batch_num, &de)) goto fail;

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
         for (ts = 0; ts < nts; ts++) {
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
               /* Temporal samples with the same sample count go to the GPU
                * together, one launch for the whole run of them. */
               nsamp = plan_nsamples(&P, batch_size);
               if (bn > 0 && (nsamp != bns || bn == MAX_BATCH_JOBS)) {
                  if (flush_batch(bjobs, bbytes, &bn, bns, P.keys + 2 * bts0, exact, opts, &bad, &sat))
                     goto devfail_de;
                  if (sat) saturated = 1;
               }
               if (bn == 0) {
                  bts0 = ts;
                  bns = nsamp;
               }
               bjobs[bn] = job;
               bbytes[bn] = job_bytes;
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
   plan_close(&P);
   return rv;
}

/* ---------------------------------------------------------------- many renders
 *
 * Small renders share the GPU: the walkers of every temporal sample of every
 * shared render run in one launch, each depositing into its own render's
 * histogram, then each render's density estimation and tone map run and all
 * images come back in one copy.  A shared render draws the walkers, keys and
 * samples fl3h_render would draw for it alone, and its integer deposits do
 * not depend on order, so its image is the one fl3h_render makes. */

/* Renders of one batch in production mode whose plotted samples stay below
 * the size from which fl3h_render bins its deposit (per-sample atomics are
 * then what fl3h_render uses too). */
static int plan_shareable(const render_plan *P, const fl3h_render_opts *opts, int bits) {
   double est;
   if (P->nbatches != 1 || P->replay || P->spec->progress) return 0;
   if (opts && opts->deposit == 2) return 0;
   if (bits && P->spec->bits != bits) return 0;
   est = P->cp.sample_density * pow(2.0, 2.0 * P->cp.zoom) * P->geo.image_width * P->geo.image_height;
   return est < (double)FL3_SHARED_MAX_SAMPLES;
}

/* No cell of a shared render can reach 2^32 in any channel, so the fast
 * deposit never saturates (fl3h_render's saturating second pass is never
 * needed for it). */
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

int fl3h_render_many(fl3h_request *reqs, int n, const fl3h_render_opts *opts) {
   render_plan *P = NULL;
   fl3_de_filters *de = NULL;
   int *slot = NULL, *sat = NULL, *widths = NULL, *heights = NULL, *live = NULL;
   double *bad = NULL;
   fl3d_multi_job *mj = NULL;
   fl3d_image_dst *dst = NULL;
   int r, ts, i, nlive = 0, nmj = 0, bits = 0, failures = 0, devok = 0;
   size_t img_total = 0, filt_total = 0;
   if (n <= 0) return 0;
   P = (render_plan *)calloc((size_t)n, sizeof(render_plan));
   de = (fl3_de_filters *)calloc((size_t)n, sizeof(fl3_de_filters));
   slot = (int *)malloc(sizeof(int) * (size_t)n);
   live = (int *)malloc(sizeof(int) * (size_t)n);
   if (!P || !de || !slot || !live) goto out;
   for (r = 0; r < n; r++) {
      reqs[r].rc = 1;
      slot[r] = -1;
   }
   /* plans and jobs of the shared renders (their frame streams advance here) */
   for (r = 0; r < n; r++) {
      fl3h_request *q = &reqs[r];
      randctx rc0;
      fl3d_multi_job *t;
      if (plan_open(&P[r], q->frame, q->out, q->field, q->nchan, q->transp, q->stats, opts)) continue;
      if (!plan_shareable(&P[r], opts, bits)) {
         plan_close(&P[r]);
         slot[r] = -2; /* rendered alone below */
         continue;
      }
      if (plan_batch_de(&P[r], 0, &de[r])) {
         plan_close(&P[r]);
         continue;
      }
      t = (fl3d_multi_job *)realloc(mj, sizeof(fl3d_multi_job) * (size_t)(nmj + P[r].nts));
      if (!t) {
         fl3_free_de_filters(&de[r]);
         plan_close(&P[r]);
         continue;
      }
      mj = t;
      rc0 = q->frame->rc;
      memset(mj + nmj, 0, sizeof(fl3d_multi_job) * (size_t)P[r].nts);
      for (ts = 0; ts < P[r].nts; ts++) {
         fl3d_multi_job *m = &mj[nmj + ts];
         double batch_size;
         void *job;
         plan_keys(&P[r], ts, ts);
         if (plan_ts_job(&P[r], ts, 1, &job, &m->bytes, &batch_size)) break;
         m->job = job;
         m->nsamples = plan_nsamples(&P[r], batch_size);
         m->key[0] = P[r].keys[2 * ts];
         m->key[1] = P[r].keys[2 * ts + 1];
         m->render = nlive;
      }
      if (ts == P[r].nts && !jobs_cannot_saturate(mj + nmj, ts)) {
         /* rendered alone after all, from the frame stream as it was */
         q->frame->rc = rc0;
         slot[r] = -2;
      }
      /* walkers per run of temporal samples with equal sample counts, as
       * fl3h_render launches them */
      if (ts == P[r].nts && slot[r] == -1)
         for (ts = 0; ts < P[r].nts;) {
            int e = ts;
            long long w;
            while (e < P[r].nts && e - ts < MAX_BATCH_JOBS && mj[nmj + e].nsamples == mj[nmj + ts].nsamples) e++;
            w = fl3d_job_walkers(mj[nmj + ts].nsamples, e - ts, opts ? opts->nwalkers : 0);
            if (w < 256) {
               fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
               break;
            }
            for (i = ts; i < e; i++) mj[nmj + i].walkers = (int)w;
            ts = e;
         }
      if (ts < P[r].nts || slot[r] == -2) {
         for (i = 0; i < P[r].nts; i++) free((void *)mj[nmj + i].job);
         fl3_free_de_filters(&de[r]);
         plan_close(&P[r]);
         continue;
      }
      nmj += P[r].nts;
      bits = P[r].spec->bits;
      slot[r] = nlive;
      live[nlive++] = r;
   }

   if (nlive > 0) {
      widths = (int *)malloc(sizeof(int) * (size_t)nlive);
      heights = (int *)malloc(sizeof(int) * (size_t)nlive);
      bad = (double *)calloc((size_t)nlive, sizeof(double));
      sat = (int *)calloc((size_t)nlive, sizeof(int));
      dst = (fl3d_image_dst *)calloc((size_t)nlive, sizeof(fl3d_image_dst));
      if (widths && heights && bad && sat && dst) {
         for (i = 0; i < nlive; i++) {
            const render_plan *Q = &P[live[i]];
            widths[i] = Q->geo.width;
            heights[i] = Q->geo.height;
            dst[i].img_off = img_total;
            dst[i].out = Q->out;
            dst[i].out_stride = plan_out_stride(Q);
            dst[i].row_bytes = (size_t)Q->geo.image_width * Q->nchan * Q->bpc;
            dst[i].rows = Q->geo.image_width > 0 ? Q->geo.image_height : 0;
            img_total += (dst[i].row_bytes * (size_t)Q->geo.image_height + 255) & ~(size_t)255;
            filt_total += (sizeof(double) * Q->filter_width * Q->filter_width + 255) & ~(size_t)255;
         }
         devok = fl3d_begin_many(widths, heights, nlive, bits) == 0 &&
                 fl3d_iterate_multi(mj, nmj, FUSE_WALKER, 0, nlive, bad, sat) == 0 &&
                 fl3d_images_begin(img_total, filt_total) == 0;
         filt_total = 0;
         for (i = 0; devok && i < nlive; i++) {
            render_plan *Q = &P[live[i]];
            fl3d_tm_params tp;
            Q->badvals = bad[i];
            devok = fl3d_select(i) == 0;
            if (devok && de[live[i]].max_filter_index == 0) {
               double k1, k2;
               plan_k1k2(Q, &k1, &k2);
               devok = fl3d_logscale(k1, k2) == 0;
            } else if (devok) {
               fl3d_de_params dp;
               plan_de_params(Q, &dp);
               devok = fl3d_density(&dp, &de[live[i]]) == 0;
            }
            plan_tm(Q, &tp);
            if (devok) devok = fl3d_tonemap_async(&tp, Q->filter, dst[i].img_off, filt_total) == 0;
            filt_total += (sizeof(double) * Q->filter_width * Q->filter_width + 255) & ~(size_t)255;
         }
         if (devok) devok = fl3d_images_fetch(dst, nlive) == 0;
         if (!devok) fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
      }
      for (i = 0; i < nlive; i++) {
         r = live[i];
         if (devok) {
            plan_finish(&P[r]);
            reqs[r].rc = 0;
         }
         fl3_free_de_filters(&de[r]);
         plan_close(&P[r]);
      }
   }
   /* the rest one at a time */
   for (r = 0; r < n; r++)
      if (slot[r] == -2)
         reqs[r].rc = fl3h_render(reqs[r].frame, reqs[r].out, reqs[r].field, reqs[r].nchan, reqs[r].transp,
                                  reqs[r].stats, opts);
out:
   for (i = 0; i < nmj; i++) free((void *)mj[i].job);
   for (r = 0; r < n; r++) failures += reqs[r].rc != 0;
   free(mj);
   free(dst);
   free(sat);
   free(bad);
   free(heights);
   free(widths);
   free(live);
   free(slot);
   free(de);
   free(P);
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
