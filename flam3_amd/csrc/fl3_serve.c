/*
 * Render server: continuous batching of independent flam3_render requests
 * (genome-search previews, animation frames) on one GPU.
 *
 * Three threads.  The planner takes the pending requests in arrival order,
 * admits as many as fit the histogram budget and the batch limit, and does
 * their host work (fl3_many_prepare: interpolation, tables, jobs, keys, DE
 * filters).  The runner executes the prepared batch on the GPU
 * (fl3_many_execute: one shared walker launch for the small renders, then
 * their density estimation and tone maps; large renders one after another).
 * The finisher copies each downloaded group of images into the callers'
 * buffers, completes the batch and frees it, while the runner already has
 * the next batch on the GPU.  While the runner has a batch on the GPU the
 * planner prepares the next one from whatever arrived meanwhile, so requests
 * that arrive while others are on the GPU are admitted as soon as the GPU
 * frees.  Every request comes out as flam3_render would make it alone.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "fl3_host.h"
#include "fl3_device.h"

typedef struct {
   fl3h_request req;
   double bytes;          /* histogram + accumulator bytes (admission) */
   int state;             /* 0 waiting, 1 admitted, 2 done */
   int rc;
   double t_submit, t_done;
} srv_rec;

/* Finisher queue entry: a group of downloaded images, or the end of a
 * batch (every group of it queued before). */
typedef struct fin_item {
   struct fin_item *next;
   fl3_images *im;            /* NULL: end of batch */
   fl3_many *B;
   fl3h_request *reqs;
   long long *tickets;
   int n;
} fin_item;

struct fl3h_server {
   pthread_mutex_t mu;
   pthread_cond_t cv_work, cv_ready, cv_done, cv_fin, cv_slot;
   pthread_t planner, runner, finisher;
   int threads;
   srv_rec **recs;        /* by ticket */
   long long nrecs, cap;
   long long next;        /* first ticket not yet admitted */
   fl3h_server_opts o;
   int has_opts;
   int stop, planner_done;
   /* the batch handed from planner to runner */
   fl3_many *ready;
   fl3h_request *ready_reqs;
   long long *ready_tickets;
   int ready_n;
   long long batches, renders;
   double prep_s, exec_s;   /* planner and runner busy time */
   /* finisher queue and the two download slots */
   fin_item *fin_head, *fin_tail;
   int runner_done;
   int slot_busy[2], next_slot;
   double dev_avail;      /* device bytes renders may use (0: unknown) */
};

/* Fraction of the device memory a batch's histograms may take: the rest is
 * for the record buffers of the binned deposit, the DE tables and images. */
#define SRV_MEM_FRACTION 0.85

static double now_s(void) {
   struct timespec t;
   clock_gettime(CLOCK_MONOTONIC, &t);
   return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Device bytes of a render's histogram and accumulator (rect.c:652-701
 * geometry, with the DE gutter bounded from above). */
static double render_bytes(const flam3_frame *f) {
   const flam3_genome *g = &f->genomes[0];
   const int ss = g->spatial_oversample > 0 ? g->spatial_oversample : 1;
   const double pad = (ceil(g->estimator) + 1.0) * ss + ss + 4.0;
   const double cells = (ss * (double)g->width + 2 * pad) * (ss * (double)g->height + 2 * pad);
   return cells * (f->bits == 64 ? 72.0 : 36.0);
}

static void *planner_main(void *arg) {
   fl3h_server *s = (fl3h_server *)arg;
   const int maxb = s->o.max_batch > 0 ? s->o.max_batch : 512;
   pthread_mutex_lock(&s->mu);
   for (;;) {
      long long t0, t1;
      double bytes = 0.0, tp, budget, mem;
      int n, i;
      fl3h_request *reqs;
      long long *tickets;
      fl3_many *B;
      while (!(s->next < s->nrecs && s->ready == NULL) && !(s->stop && s->next == s->nrecs))
         pthread_cond_wait(&s->cv_work, &s->mu);
      if (s->next == s->nrecs && s->stop) break;
      /* Admission by histogram memory: a batch's histograms and accumulators
       * fit the device memory the renders may use (refreshed after every
       * batch), and by default one shared launch's cell space (32e6 cells):
       * the binned deposit bins a shared group of at most that many cells,
       * and groups of a batch run one after another on the same buffers.  A
       * request that alone exceeds the device memory fails without running. */
      mem = s->dev_avail > 0 ? SRV_MEM_FRACTION * s->dev_avail : 0.0;
      budget = s->o.hist_budget_bytes > 0 ? s->o.hist_budget_bytes : 32e6 * 36.0;
      if (mem > 0 && budget > mem) budget = mem;
      while (s->next < s->nrecs && mem > 0 && s->recs[s->next]->bytes > mem) {
         srv_rec *r = s->recs[s->next++];
         fprintf(stderr, "flam3_amd: render server: a render needing %.3g GB of histograms exceeds the %.3g GB "
                         "of device memory available\n", r->bytes / 1e9, mem / 1e9);
         r->rc = 1;
         r->state = 2;
         r->t_done = now_s();
         pthread_cond_broadcast(&s->cv_done);
      }
      if (s->next == s->nrecs) continue;
      /* admission in arrival order: the batch limit and the budget (the
       * first request always goes) */
      t0 = s->next;
      for (t1 = t0; t1 < s->nrecs && t1 - t0 < maxb; t1++) {
         if (t1 > t0 && bytes + s->recs[t1]->bytes > budget) break;
         if (mem > 0 && s->recs[t1]->bytes > mem) break;
         bytes += s->recs[t1]->bytes;
         s->recs[t1]->state = 1;
      }
      s->next = t1;
      n = (int)(t1 - t0);
      reqs = (fl3h_request *)malloc(sizeof(fl3h_request) * (size_t)n);
      tickets = (long long *)malloc(sizeof(long long) * (size_t)n);
      if (reqs && tickets)
         for (i = 0; i < n; i++) {
            reqs[i] = s->recs[t0 + i]->req;
            tickets[i] = t0 + i;
         }
      pthread_mutex_unlock(&s->mu);
      tp = now_s();
      B = reqs && tickets ? fl3_many_prepare(reqs, n, s->has_opts ? &s->o.opts : NULL) : NULL;
      tp = now_s() - tp;
      pthread_mutex_lock(&s->mu);
      s->prep_s += tp;
      if (!B) {
         for (i = 0; i < n; i++) {
            srv_rec *r = s->recs[t0 + i];
            r->rc = 1;
            r->state = 2;
            r->t_done = now_s();
         }
         free(reqs);
         free(tickets);
         pthread_cond_broadcast(&s->cv_done);
         continue;
      }
      s->ready = B;
      s->ready_reqs = reqs;
      s->ready_tickets = tickets;
      s->ready_n = n;
      pthread_cond_signal(&s->cv_ready);
   }
   s->planner_done = 1;
   pthread_cond_signal(&s->cv_ready);
   pthread_mutex_unlock(&s->mu);
   return NULL;
}

static void fin_push(fl3h_server *s, fin_item *it) {
   pthread_mutex_lock(&s->mu);
   it->next = NULL;
   if (s->fin_tail) s->fin_tail->next = it;
   else s->fin_head = it;
   s->fin_tail = it;
   pthread_cond_signal(&s->cv_fin);
   pthread_mutex_unlock(&s->mu);
}

/* fl3_delivery hooks (runner thread) */
static int slot_acquire(void *ctx) {
   fl3h_server *s = (fl3h_server *)ctx;
   int k;
   pthread_mutex_lock(&s->mu);
   while (s->slot_busy[s->next_slot]) pthread_cond_wait(&s->cv_slot, &s->mu);
   k = s->next_slot;
   s->slot_busy[k] = 1;
   s->next_slot ^= 1;
   pthread_mutex_unlock(&s->mu);
   return k;
}

static void images_hand_over(void *ctx, fl3_images *im) {
   fin_item *it = (fin_item *)calloc(1, sizeof(fin_item));
   if (!it) {   /* deliver here rather than lose the group */
      fl3h_server *s = (fl3h_server *)ctx;
      const int k = fl3_images_deliver(im);
      pthread_mutex_lock(&s->mu);
      if (k >= 0) s->slot_busy[k] = 0;
      pthread_cond_broadcast(&s->cv_slot);
      pthread_mutex_unlock(&s->mu);
      return;
   }
   it->im = im;
   fin_push((fl3h_server *)ctx, it);
}

static void *finisher_main(void *arg) {
   fl3h_server *s = (fl3h_server *)arg;
   pthread_mutex_lock(&s->mu);
   for (;;) {
      fin_item *it;
      int i;
      while (s->fin_head == NULL && !s->runner_done) pthread_cond_wait(&s->cv_fin, &s->mu);
      if (s->fin_head == NULL) break;
      it = s->fin_head;
      s->fin_head = it->next;
      if (!s->fin_head) s->fin_tail = NULL;
      pthread_mutex_unlock(&s->mu);
      if (it->im) {
         const int k = fl3_images_deliver(it->im);
         pthread_mutex_lock(&s->mu);
         if (k >= 0) s->slot_busy[k] = 0;
         pthread_cond_broadcast(&s->cv_slot);
      } else {
         fl3_many_free(it->B);
         pthread_mutex_lock(&s->mu);
         for (i = 0; i < it->n; i++) {
            srv_rec *r = s->recs[it->tickets[i]];
            r->rc = it->reqs[i].rc;
            r->state = 2;
            r->t_done = now_s();
         }
         s->batches++;
         s->renders += it->n;
         free(it->reqs);
         free(it->tickets);
         pthread_cond_broadcast(&s->cv_done);
      }
      free(it);
   }
   pthread_mutex_unlock(&s->mu);
   return NULL;
}

static void *runner_main(void *arg) {
   fl3h_server *s = (fl3h_server *)arg;
   pthread_mutex_lock(&s->mu);
   for (;;) {
      fl3_many *B;
      fl3h_request *reqs;
      long long *tickets;
      int n, i;
      double te;
      while (s->ready == NULL && !s->planner_done) pthread_cond_wait(&s->cv_ready, &s->mu);
      if (s->ready == NULL) break;
      B = s->ready;
      reqs = s->ready_reqs;
      tickets = s->ready_tickets;
      n = s->ready_n;
      s->ready = NULL;
      pthread_cond_signal(&s->cv_work);   /* the planner may prepare the next batch */
      pthread_mutex_unlock(&s->mu);
      te = now_s();
      {
         fl3_delivery dv;
         fin_item *end = (fin_item *)calloc(1, sizeof(fin_item));
         dv.ctx = s;
         dv.acquire = slot_acquire;
         dv.hand_over = images_hand_over;
         if (end) fl3_many_set_delivery(B, &dv);   /* else the images are copied in place */
         fl3_many_execute(B);
         te = now_s() - te;
         {
            double avail = 0.0;
            if (fl3d_device_memory(&avail, NULL) == 0) {
               pthread_mutex_lock(&s->mu);
               s->dev_avail = avail;
               pthread_mutex_unlock(&s->mu);
            }
         }
         if (end) {
            /* the finisher completes the batch after its image groups */
            end->B = B;
            end->reqs = reqs;
            end->tickets = tickets;
            end->n = n;
            fin_push(s, end);
            pthread_mutex_lock(&s->mu);
            s->exec_s += te;
            continue;
         }
      }
      fl3_many_free(B);
      pthread_mutex_lock(&s->mu);
      s->exec_s += te;
      for (i = 0; i < n; i++) {
         srv_rec *r = s->recs[tickets[i]];
         r->rc = reqs[i].rc;
         r->state = 2;
         r->t_done = now_s();
      }
      s->batches++;
      s->renders += n;
      free(reqs);
      free(tickets);
      pthread_cond_broadcast(&s->cv_done);
   }
   s->runner_done = 1;
   pthread_cond_signal(&s->cv_fin);
   pthread_mutex_unlock(&s->mu);
   return NULL;
}

/* Stop and join the threads started so far: the planner ends when every
 * admitted request is prepared, the runner when the planner has ended and
 * the last batch ran, the finisher when the runner has ended and its queue
 * is empty.  A thread that was never started is marked ended here. */
static void stop_threads(fl3h_server *s) {
   pthread_mutex_lock(&s->mu);
   s->stop = 1;
   pthread_cond_broadcast(&s->cv_work);
   if (s->threads < 2) s->planner_done = 1;
   if (s->threads < 3) s->runner_done = 1;
   pthread_cond_broadcast(&s->cv_ready);
   pthread_cond_broadcast(&s->cv_fin);
   pthread_mutex_unlock(&s->mu);
   if (s->threads >= 2) pthread_join(s->planner, NULL);
   if (s->threads >= 3) pthread_join(s->runner, NULL);
   if (s->threads >= 1) pthread_join(s->finisher, NULL);
}

fl3h_server *fl3h_server_create(const fl3h_server_opts *o) {
   fl3h_server *s = (fl3h_server *)calloc(1, sizeof(fl3h_server));
   if (!s) return NULL;
   if (o) {
      s->o = *o;
      s->has_opts = 1;
   }
   pthread_mutex_init(&s->mu, NULL);
   pthread_cond_init(&s->cv_work, NULL);
   pthread_cond_init(&s->cv_ready, NULL);
   pthread_cond_init(&s->cv_done, NULL);
   pthread_cond_init(&s->cv_fin, NULL);
   pthread_cond_init(&s->cv_slot, NULL);
   if (fl3d_device_memory(&s->dev_avail, NULL)) s->dev_avail = 0.0;
   {
      /* the device buffers of a full batch, allocated now: growing them
       * between batches stalls the device for milliseconds (an unlucky
       * batch's requests then wait tens of ms); if it fails, they grow as
       * needed */
      double budget = s->o.hist_budget_bytes > 0 ? s->o.hist_budget_bytes : 32e6 * 36.0;
      if (s->dev_avail > 0 && budget > SRV_MEM_FRACTION * s->dev_avail) budget = SRV_MEM_FRACTION * s->dev_avail;
      (void)fl3d_reserve(budget);
   }
   if (pthread_create(&s->finisher, NULL, finisher_main, s)) goto bad;
   s->threads = 1;
   if (pthread_create(&s->planner, NULL, planner_main, s)) {
      stop_threads(s);
      goto bad;
   }
   s->threads = 2;
   if (pthread_create(&s->runner, NULL, runner_main, s)) {
      stop_threads(s);
      goto bad;
   }
   s->threads = 3;
   return s;
bad:
   pthread_cond_destroy(&s->cv_slot);
   pthread_cond_destroy(&s->cv_fin);
   pthread_cond_destroy(&s->cv_done);
   pthread_cond_destroy(&s->cv_ready);
   pthread_cond_destroy(&s->cv_work);
   pthread_mutex_destroy(&s->mu);
   free(s);
   return NULL;
}

long long fl3h_server_submit(fl3h_server *s, flam3_frame *f, void *out, int field, int nchan, int transp,
                             stat_struct *stats) {
   srv_rec *r;
   long long t;
   if (!s || !f || !out || !stats) return -1;
   r = (srv_rec *)calloc(1, sizeof(srv_rec));
   if (!r) return -1;
   r->req.frame = f;
   r->req.out = out;
   r->req.field = field;
   r->req.nchan = nchan;
   r->req.transp = transp;
   r->req.stats = stats;
   r->req.rc = 1;
   r->bytes = f->genomes && f->ngenomes > 0 ? render_bytes(f) : 0.0;
   pthread_mutex_lock(&s->mu);
   if (s->stop) {
      pthread_mutex_unlock(&s->mu);
      free(r);
      return -1;
   }
   if (s->nrecs == s->cap) {
      long long nc = s->cap ? 2 * s->cap : 1024;
      srv_rec **nr = (srv_rec **)realloc(s->recs, sizeof(srv_rec *) * (size_t)nc);
      if (!nr) {
         pthread_mutex_unlock(&s->mu);
         free(r);
         return -1;
      }
      s->recs = nr;
      s->cap = nc;
   }
   r->t_submit = now_s();
   t = s->nrecs++;
   s->recs[t] = r;
   pthread_cond_signal(&s->cv_work);
   pthread_mutex_unlock(&s->mu);
   return t;
}

int fl3h_server_wait(fl3h_server *s, long long ticket, double *latency_seconds) {
   int rc;
   if (!s) return 1;
   pthread_mutex_lock(&s->mu);
   if (ticket < 0 || ticket >= s->nrecs) {
      pthread_mutex_unlock(&s->mu);
      return 1;
   }
   while (s->recs[ticket]->state != 2) pthread_cond_wait(&s->cv_done, &s->mu);
   rc = s->recs[ticket]->rc;
   if (latency_seconds) *latency_seconds = s->recs[ticket]->t_done - s->recs[ticket]->t_submit;
   pthread_mutex_unlock(&s->mu);
   return rc;
}

int fl3h_server_done(fl3h_server *s, long long ticket) {
   int d;
   if (!s) return 0;
   pthread_mutex_lock(&s->mu);
   d = ticket >= 0 && ticket < s->nrecs && s->recs[ticket]->state == 2;
   pthread_mutex_unlock(&s->mu);
   return d;
}

void fl3h_server_counts(fl3h_server *s, long long *batches, long long *renders, double *prep_seconds,
                        double *gpu_seconds) {
   if (!s) return;
   pthread_mutex_lock(&s->mu);
   if (batches) *batches = s->batches;
   if (renders) *renders = s->renders;
   if (prep_seconds) *prep_seconds = s->prep_s;
   if (gpu_seconds) *gpu_seconds = s->exec_s;
   pthread_mutex_unlock(&s->mu);
}

void fl3h_server_destroy(fl3h_server *s) {
   long long i;
   if (!s) return;
   stop_threads(s);
   for (i = 0; i < s->nrecs; i++) free(s->recs[i]);
   free(s->recs);
   pthread_cond_destroy(&s->cv_slot);
   pthread_cond_destroy(&s->cv_fin);
   pthread_cond_destroy(&s->cv_done);
   pthread_cond_destroy(&s->cv_ready);
   pthread_cond_destroy(&s->cv_work);
   pthread_mutex_destroy(&s->mu);
   free(s);
}
