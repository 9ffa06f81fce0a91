/*
 * Genome ingestion on a host thread pool: XML documents parsed by the
 * caller's flam3 parser (flam3_parse_xml2, parser.c:220-1256 via
 * flam3.c:1423-1561) on several threads, ahead of the renders that use them
 * (a render server's request loop, a preview loop).
 *
 * The parser keeps lazily initialised process state (the XML library, the
 * palette library read on first use), so the first document is parsed alone
 * and the pool parses in parallel only after it has returned.  Each parsed
 * genome is checked for what the renderer relies on and the parser
 * establishes (parser.c:802-821: the final xform last; sizes, oversampling,
 * batches, temporal samples and quality positive; xaos rows present for the
 * standard xforms).  Results come back per ticket, in any order of
 * completion.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"

typedef struct {
   char *xml, *name;
   flam3_genome *genomes;
   int ncps, state, rc;   /* state: 0 queued, 1 parsing, 2 done */
} ing_job;

struct fl3h_ingest {
   pthread_mutex_t mu;
   pthread_cond_t cv_work, cv_done;
   pthread_t *threads;
   int nthreads, started;
   fl3h_parse_fn parse;
   int default_flag;
   ing_job **jobs;
   long long njobs, cap, next;
   int warm;              /* 0 none parsed yet, 1 first parse running, 2 warmed up */
   int stop;
};

/* What the renderer relies on in a parsed genome; 0 if it holds. */
static int check_genome(const flam3_genome *g, char *why, size_t n) {
   int nstd;
   if (g->num_xforms < 0 || (g->num_xforms > 0 && !g->xform)) {
      snprintf(why, n, "no xform storage");
      return 1;
   }
   if (g->final_xform_index >= 0 &&
       (g->final_xform_index != g->num_xforms - 1 || !g->final_xform_enable)) {
      snprintf(why, n, "final xform %d is not the last of %d xforms", g->final_xform_index, g->num_xforms);
      return 1;
   }
   if (g->final_xform_index < 0 && g->final_xform_enable) {
      snprintf(why, n, "final xform enabled without one");
      return 1;
   }
   nstd = g->num_xforms - (g->final_xform_index >= 0 ? 1 : 0);
   if (nstd > 0 && !g->chaos) {
      snprintf(why, n, "no xaos rows for %d xforms", nstd);
      return 1;
   }
   if (g->width <= 0 || g->height <= 0 || g->spatial_oversample < 1 || g->nbatches < 1 ||
       g->ntemporal_samples < 1 || !(g->sample_density > 0.0)) {
      snprintf(why, n, "size %dx%d, oversample %d, batches %d, temporal samples %d, quality %g", g->width,
               g->height, g->spatial_oversample, g->nbatches, g->ntemporal_samples, g->sample_density);
      return 1;
   }
   return 0;
}

static void *worker_main(void *arg) {
   fl3h_ingest *q = (fl3h_ingest *)arg;
   pthread_mutex_lock(&q->mu);
   for (;;) {
      ing_job *j;
      int ncps = 0, rc = 0, first, i;
      flam3_genome *gs;
      char why[160];
      while (q->warm == 1 || (q->next == q->njobs && !q->stop)) pthread_cond_wait(&q->cv_work, &q->mu);
      if (q->next == q->njobs) break;   /* stopping and nothing queued */
      j = q->jobs[q->next++];
      j->state = 1;
      first = q->warm == 0;
      if (first) q->warm = 1;
      pthread_mutex_unlock(&q->mu);
      gs = q->parse(j->xml, j->name, q->default_flag, &ncps);
      if (!gs || ncps < 1) {
         fprintf(stderr, "flam3_amd: ingest: %s: the parser returned no genome\n", j->name);
         rc = 1;
      }
      for (i = 0; !rc && i < ncps; i++)
         if (check_genome(&gs[i], why, sizeof(why))) {
            fprintf(stderr, "flam3_amd: ingest: %s: genome %d: %s\n", j->name, i, why);
            rc = 1;
         }
      pthread_mutex_lock(&q->mu);
      j->genomes = gs;
      j->ncps = gs ? ncps : 0;
      j->rc = rc;
      j->state = 2;
      free(j->xml);
      j->xml = NULL;
      if (first) {
         q->warm = 2;
         pthread_cond_broadcast(&q->cv_work);
      }
      pthread_cond_broadcast(&q->cv_done);
   }
   pthread_mutex_unlock(&q->mu);
   return NULL;
}

fl3h_ingest *fl3h_ingest_create(fl3h_parse_fn parse, int nthreads, int default_flag) {
   fl3h_ingest *q;
   int i;
   if (!parse) return NULL;
   if (nthreads < 1) nthreads = 8;
   q = (fl3h_ingest *)calloc(1, sizeof(fl3h_ingest));
   if (!q) return NULL;
   q->parse = parse;
   q->default_flag = default_flag;
   q->threads = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
   if (!q->threads) {
      free(q);
      return NULL;
   }
   pthread_mutex_init(&q->mu, NULL);
   pthread_cond_init(&q->cv_work, NULL);
   pthread_cond_init(&q->cv_done, NULL);
   for (i = 0; i < nthreads; i++) {
      if (pthread_create(&q->threads[i], NULL, worker_main, q)) break;
      q->started++;
   }
   q->nthreads = nthreads;
   if (q->started == 0) {
      fl3h_ingest_destroy(q);
      return NULL;
   }
   return q;
}

long long fl3h_ingest_submit(fl3h_ingest *q, const char *xml, const char *filename) {
   ing_job *j;
   long long t;
   if (!q || !xml) return -1;
   j = (ing_job *)calloc(1, sizeof(ing_job));
   if (!j) return -1;
   j->xml = strdup(xml);
   j->name = strdup(filename ? filename : "(ingest)");
   if (!j->xml || !j->name) {
      free(j->xml);
      free(j->name);
      free(j);
      return -1;
   }
   pthread_mutex_lock(&q->mu);
   if (q->stop) {
      pthread_mutex_unlock(&q->mu);
      free(j->xml);
      free(j->name);
      free(j);
      return -1;
   }
   if (q->njobs == q->cap) {
      const long long nc = q->cap ? 2 * q->cap : 256;
      ing_job **nj = (ing_job **)realloc(q->jobs, sizeof(ing_job *) * (size_t)nc);
      if (!nj) {
         pthread_mutex_unlock(&q->mu);
         free(j->xml);
         free(j->name);
         free(j);
         return -1;
      }
      q->jobs = nj;
      q->cap = nc;
   }
   t = q->njobs++;
   q->jobs[t] = j;
   pthread_cond_signal(&q->cv_work);
   pthread_mutex_unlock(&q->mu);
   return t;
}

int fl3h_ingest_wait(fl3h_ingest *q, long long ticket, flam3_genome **genomes, int *ncps) {
   ing_job *j;
   int rc;
   if (!q || ticket < 0) return 1;
   pthread_mutex_lock(&q->mu);
   if (ticket >= q->njobs || !q->jobs[ticket]) {
      pthread_mutex_unlock(&q->mu);
      return 1;
   }
   j = q->jobs[ticket];
   while (j->state != 2) pthread_cond_wait(&q->cv_done, &q->mu);
   q->jobs[ticket] = NULL;   /* handed over */
   pthread_mutex_unlock(&q->mu);
   if (genomes) *genomes = j->genomes;
   if (ncps) *ncps = j->ncps;
   rc = j->rc;
   free(j->name);
   free(j);
   return rc;
}

void fl3h_ingest_destroy(fl3h_ingest *q) {
   long long t;
   int i;
   if (!q) return;
   pthread_mutex_lock(&q->mu);
   q->stop = 1;
   pthread_cond_broadcast(&q->cv_work);
   pthread_mutex_unlock(&q->mu);
   for (i = 0; i < q->started; i++) pthread_join(q->threads[i], NULL);
   /* results nobody waited for: the genomes belong to the parser's
    * allocator, so they are left to the caller's process (as after a
    * flam3_parse_xml2 whose result is dropped) */
   for (t = 0; t < q->njobs; t++)
      if (q->jobs[t]) {
         free(q->jobs[t]->xml);
         free(q->jobs[t]->name);
         free(q->jobs[t]);
      }
   free(q->jobs);
   free(q->threads);
   pthread_cond_destroy(&q->cv_done);
   pthread_cond_destroy(&q->cv_work);
   pthread_mutex_destroy(&q->mu);
   free(q);
}
