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

#define WHITE_LEVEL 255.0
#define PREFILTER_WHITE 255
#define FUSE_NORMAL 15   /* rect.c:36 */
#define FUSE_WALKER 100
#define FUSE_EARLYCLIP 100

/* Job records carry the job index in 16 bits. */
#define MAX_BATCH_JOBS 4096

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

int fl3h_render(flam3_frame *spec, void *out, int field, int nchan, int transp, stat_struct *stats,
                const fl3h_render_opts *opts) {
   flam3_genome cp;
   geometry geo;
   fl3h_camera *cam = NULL;
   double *filter = NULL, *temporal_filter = NULL, *temporal_deltas = NULL;
   double sumfilt, vibrancy = 0.0, gamma = 0.0, background[3] = {0.0, 0.0, 0.0};
   double badvals = 0.0;
   int i, j, batch_num, ts, nbatches, nts, oversample, filter_width, max_de_fw, vib_gam_n = 0;
   int cmap_size, nthreads, rv = 1, aborted = 0;
   const int replay = opts && opts->replay;
   const int bpc = spec->bytes_per_channel;
   uint32_t *walker_states = NULL;
   uint64_t *keys = NULL;
   void **bjobs = NULL;   /* temporal samples waiting for their launch */
   size_t *bbytes = NULL;
   int bn = 0, bts0 = 0;
   int64_t bns = 0;
   time_t tstart = time(NULL);

   memset(&cp, 0, sizeof(cp));
   if (spec->bits != 32 && spec->bits != 33 && spec->bits != 64) {
      fprintf(stderr, "flam3: bits must be 32, 33, or 64 not %d.\n", spec->bits);
      return 1;
   }
   nthreads = spec->nthreads > 0 ? spec->nthreads : 1;
   stats->num_iters = 0;
   /* apophysis palettes use 255 entries (rect.c:587) */
   cmap_size = 256 - env_int("apo_palette", 0);

   if (fl3_interpolate(spec->genomes, spec->ngenomes, spec->time, 0, &cp)) return 1;
   oversample = cp.spatial_oversample;
   nbatches = cp.nbatches;
   nts = cp.ntemporal_samples;
   if (nbatches < 1) {
      fprintf(stderr, "nbatches must be positive, not %d.\n", nbatches);
      goto fail;
   }
   if (oversample < 1) {
      fprintf(stderr, "oversample must be positive, not %d.\n", oversample);
      goto fail;
   }

   memset(&geo, 0, sizeof(geo));
   geo.image_width = cp.width;
   geo.out_width = cp.width;
   if (field) {
      geo.image_height = cp.height / 2;
      if (field == flam3_field_odd) out = (unsigned char *)out + nchan * bpc * geo.out_width;
      geo.out_width *= 2;
   } else
      geo.image_height = cp.height;

   filter_width = fl3_create_spatial_filter(&spec->genomes[0], spec->pixel_aspect_ratio, field, &filter);
   if (filter_width < 0) {
      fprintf(stderr, "flam3_create_spatial_filter returned error: aborting\n");
      goto fail;
   }
   sumfilt = fl3_create_temporal_filter(nbatches * nts, cp.temporal_filter_type, cp.temporal_filter_exp,
                                        cp.temporal_filter_width, &temporal_filter, &temporal_deltas);

   /* gutter for the spatial filter, widened for the DE radius (rect.c:652-681) */
   geo.gutter = (filter_width - oversample) / 2;
   max_de_fw = 0;
   for (i = 0; i < spec->ngenomes; i++) {
      int w = (int)ceil(spec->genomes[i].estimator) * oversample;
      if (w > max_de_fw) max_de_fw = w;
   }
   if (max_de_fw > 0) max_de_fw += oversample - 1;
   if (max_de_fw > geo.gutter) {
      geo.de_offset = max_de_fw - geo.gutter;
      geo.gutter = max_de_fw;
   } else
      geo.de_offset = 0;
   geo.height = oversample * geo.image_height + 2 * geo.gutter;
   geo.width = oversample * geo.image_width + 2 * geo.gutter;

   if (fl3d_begin(geo.width, geo.height, spec->bits)) {
      fprintf(stderr, "flam3_amd: %s\n", fl3d_error());
      goto fail;
   }
   cam = (fl3h_camera *)calloc(1, sizeof(fl3h_camera));
   walker_states = (uint32_t *)malloc(sizeof(uint32_t) * 36 * (size_t)nthreads);
   keys = (uint64_t *)calloc((size_t)2 * nts, sizeof(uint64_t));
   bjobs = (void **)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(void *));
   bbytes = (size_t *)calloc((size_t)(nts < MAX_BATCH_JOBS ? nts : MAX_BATCH_JOBS), sizeof(size_t));
   if (!cam || !walker_states || !keys || !bjobs || !bbytes) goto fail;

   for (batch_num = 0; batch_num < nbatches && !aborted; batch_num++) {
      fl3_de_filters de;
      double k1, area, k2, batch_bad = 0.0;
      int saturated = 0, exact = 0, pass;
      memset(&de, 0, sizeof(de));

      /* genome at the batch's first temporal sample, for the DE widths */
      if (fl3_interpolate(spec->genomes, spec->ngenomes, spec->time + temporal_deltas[batch_num * nts], 0, &cp))
         goto fail;
      if (cp.estimator < 0.0 || cp.estimator_minimum < 0.0) {
         fprintf(stderr, "density estimator filter widths must be >= 0\n");
         goto fail;
      }
      if (spec->bits <= 32 && cp.estimator > 0.0) {
         fprintf(stderr, "warning: density estimation disabled with %d bit buffers.\n", spec->bits);
         cp.estimator = 0.0;
      }
      if (cp.estimator > 0.0) {
         if (fl3_create_de_filters(cp.estimator, cp.estimator_minimum, cp.estimator_curve, oversample, &de)) {
            fprintf(stderr, "de.kernel_size returned 0 - aborting.\n");
            goto fail;
         }
      } else
         de.max_filter_index = 0;

      /* Pass 0 uses fast deposits; if a hit counter saturated, pass 1
       * repeats the batch with saturating adds on every channel, reusing
       * the walker keys so the same points are drawn. */
      for (pass = 0; pass < 2; pass++) {
         if (pass == 1) {
            if (!saturated || replay) break;
            exact = 1;
            badvals -= batch_bad;
            batch_bad = 0.0;
            if (fl3d_clear_buckets()) goto devfail;
         }
         for (ts = 0; ts < nts; ts++) {
            const int tidx = batch_num * nts + ts;
            const double color_scalar = temporal_filter[tidx];
            unsigned short *distrib;
            void *job;
            size_t job_bytes;
            double batch_size, bad = 0.0;
            int sat = 0, nrows;

            if (fl3_interpolate(spec->genomes, spec->ngenomes, spec->time + temporal_deltas[tidx], 0, &cp))
               goto fail;
            if (fl3_prepare_precalc_flags(&cp)) {
               fprintf(stderr, "prepare xform pointers returned error: aborting.\n");
               goto fail;
            }
            for (i = 0; i < cp.num_xforms; i++) fl3_xform_precalc(&cp.xform[i]);
            distrib = fl3_create_xform_distrib(&cp, &nrows);
            if (!distrib) {
               fprintf(stderr, "create xform distrib returned error: aborting.\n");
               goto fail;
            }
            /* palette scaled to the white level and the temporal weight (rect.c:778-782) */
            for (j = 0; j < 256; j++)
               for (i = 0; i < 4; i++) cam->dmap[j][i] = cp.palette[j].color[i] * WHITE_LEVEL * color_scalar;
            cam->cmap_size = cmap_size;
            cam->palette_mode = cp.palette_mode;
            if (setup_camera(spec, &cp, field, oversample, &geo, cam)) {
               free(distrib);
               goto fail;
            }
            batch_size = geo.sample_density * geo.image_width * geo.image_height / (nbatches * nts);
            if (pass == 0) stats->num_iters += batch_size;

            job = fl3_pack_job(&cp, distrib, cam, &job_bytes);
            free(distrib);
            if (!job) goto fail;

            if (poll_progress(spec, 100.0 * (((double)ts / nts) + batch_num) / nbatches, 0, 0.0)) {
               free(job);
               aborted = 1;
               break;
            }

            if (replay) {
               /* walker w = reference thread w (rect.c:858-865) */
               for (i = 0; i < nthreads; i++) thread_state(&spec->rc, walker_states + 36 * i);
               if (fl3d_replay(job, job_bytes, walker_states, nthreads, batch_size / (double)nthreads,
                               spec->sub_batch_size > 0 ? spec->sub_batch_size : 10000,
                               spec->earlyclip ? FUSE_EARLYCLIP : FUSE_NORMAL, &bad)) {
                  free(job);
                  goto devfail;
               }
            } else {
               int64_t nsamp;
               if (pass == 0) {
                  /* The frame stream advances by 16 words per reference thread,
                   * as rect.c:858-865 consumes it; the first four seed the
                   * walker key. */
                  uint64_t w[4];
                  for (i = 0; i < nthreads; i++)
                     for (j = 0; j < FLAM3_RANDSIZ; j++) {
                        ub4 v = fl3_irand(&spec->rc);
                        if (i == 0 && j < 4) w[j] = (uint64_t)(v & 0xffffffffUL);
                     }
                  keys[2 * ts] = w[0] | (w[1] << 32);
                  keys[2 * ts + 1] = (w[2] | (w[3] << 32)) ^ ((uint64_t)tidx * 0x9E3779B97F4A7C15ULL);
               }
               /* Temporal samples with the same sample count go to the GPU
                * together, one launch for the whole run of them. */
               nsamp = (int64_t)nthreads * (int64_t)(batch_size / (double)nthreads);
               if (bn > 0 && (nsamp != bns || bn == MAX_BATCH_JOBS)) {
                  if (flush_batch(bjobs, bbytes, &bn, bns, keys + 2 * bts0, exact, opts, &bad, &sat)) goto devfail;
                  badvals += bad;
                  batch_bad += bad;
                  if (sat) saturated = 1;
                  bad = 0.0;
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
            badvals += bad;
            batch_bad += bad;
            if (pass == 0) {
               vibrancy += cp.vibrancy;
               gamma += cp.gamma;
               background[0] += cp.background[0];
               background[1] += cp.background[1];
               background[2] += cp.background[2];
               vib_gam_n++;
            }
         }
         if (!aborted && bn > 0) {
            double bad = 0.0;
            int sat = 0;
            if (flush_batch(bjobs, bbytes, &bn, bns, keys + 2 * bts0, exact, opts, &bad, &sat)) goto devfail;
            badvals += bad;
            batch_bad += bad;
            if (sat) saturated = 1;
         }
         if (aborted) break;
      }
      if (aborted) {
         fl3_free_de_filters(&de);
         break;
      }

      /* density normalisation of this batch (rect.c:933-937) */
      k1 = (cp.contrast * cp.brightness * PREFILTER_WHITE * 268.0 * (1.0 / (double)nbatches)) / 256;
      area = geo.image_width * geo.image_height / (geo.ppux * geo.ppuy);
      k2 = (oversample * oversample * nbatches) / (cp.contrast * area * WHITE_LEVEL * geo.sample_density * sumfilt);

      if (de.max_filter_index == 0) {
         if (fl3d_logscale(k1, k2)) goto devfail;
      } else {
         fl3d_de_params dp;
         dp.oversample = oversample;
         dp.k1 = k1;
         dp.k2 = k2;
         dp.curve = cp.estimator_curve;
         dp.exact = replay || spec->bits == 64;
         if (fl3d_density(&dp, &de)) {
            fl3_free_de_filters(&de);
            goto devfail;
         }
         if (poll_progress(spec, 100.0, 1, 0.0)) aborted = 1;
      }
      fl3_free_de_filters(&de);
   }

   if (!aborted) {
      fl3d_tm_params tp;
      memset(&tp, 0, sizeof(tp));
      vibrancy /= vib_gam_n;
      background[0] /= vib_gam_n / 256.0;
      background[1] /= vib_gam_n / 256.0;
      background[2] /= vib_gam_n / 256.0;
      tp.image_width = geo.image_width;
      tp.image_height = geo.image_height;
      tp.oversample = oversample;
      tp.de_offset = geo.de_offset;
      tp.filter_width = filter_width;
      tp.nchan = nchan;
      tp.transp = transp;
      tp.bytes_per_channel = bpc;
      tp.earlyclip = spec->earlyclip;
      tp.g = 1.0 / (gamma / vib_gam_n);
      tp.linrange = cp.gam_lin_thresh;
      tp.vibrancy = vibrancy;
      tp.highpow = cp.highlight_power;
      for (i = 0; i < 3; i++) tp.background[i] = background[i];
      if (fl3d_tonemap(&tp, filter, out, (size_t)nchan * bpc * geo.out_width)) goto devfail;
   }

   stats->badvals = badvals;
   if (!aborted && getenv("insert_palette")) {
      /* palette strip along the top (rect.c:1267-1279) */
      int ph = 100;
      if (ph >= geo.image_height) ph = geo.image_height;
      for (j = 0; j < ph; j++)
         for (i = 0; i < geo.image_width; i++) {
            unsigned char *p = (unsigned char *)out + nchan * (i + j * geo.out_width);
            p[0] = (unsigned char)cam->dmap[i * 256 / geo.image_width][0];
            p[1] = (unsigned char)cam->dmap[i * 256 / geo.image_width][1];
            p[2] = (unsigned char)cam->dmap[i * 256 / geo.image_width][2];
         }
   }
   stats->render_seconds = (int)(time(NULL) - tstart);
   rv = 0;
   goto cleanup;

devfail:
   fprintf(stderr, "flam3_amd: device error: %s\n", fl3d_error());
fail:
   rv = 1;
cleanup:
   free(filter);
   free(temporal_filter);
   free(temporal_deltas);
   free(cam);
   free(walker_states);
   free(keys);
   for (i = 0; i < bn; i++) free(bjobs[i]);
   free(bjobs);
   free(bbytes);
   fl3_genome_free(&cp);
   return rv;
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
