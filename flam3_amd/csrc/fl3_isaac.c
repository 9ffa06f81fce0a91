/*
 * ISAAC, Bob Jenkins' generator, at the reference's size (RANDSIZL = 4,
 * isaac.h:19) and with the reference's LP64 word type: ub4 is
 * `unsigned long` (isaacs.h:17), so the seeding scramble runs on 64-bit
 * words and its right shifts pull upper bits into the low 32 (isaac.c:48-58).
 * The refill masks every result to 32 bits (isaac.c:17-23), so only the
 * seeding needs the 64-bit words.  These steps have to match the reference
 * exactly: the walker streams and the caller-visible flam3_frame.rc advance
 * through them.
 */
#include "fl3_host.h"

#define M32 0xffffffffUL

/* One refill of 16 outputs (isaac.c:25-45). */
void fl3_isaac_refill(randctx *ctx) {
   ub4 *mm = ctx->randmem, *rr = ctx->randrsl;
   ub4 a = ctx->randa;
   ub4 b = (ctx->randb + (++ctx->randc)) & M32;
   int i;
   for (i = 0; i < FLAM3_RANDSIZ; i++) {
      ub4 x = mm[i], y, mixv;
      switch (i & 3) {
         case 0: mixv = a << 13; break;
         case 1: mixv = a >> 6; break;
         case 2: mixv = a << 2; break;
         default: mixv = a >> 16; break;
      }
      a = ((a ^ mixv) + mm[(i + FLAM3_RANDSIZ / 2) & (FLAM3_RANDSIZ - 1)]) & M32;
      y = (mm[(x >> 2) & (FLAM3_RANDSIZ - 1)] + a + b) & M32;
      mm[i] = y;
      b = (mm[((y >> FLAM3_RANDSIZL) >> 2) & (FLAM3_RANDSIZ - 1)] + x) & M32;
      rr[i] = b;
   }
   ctx->randb = b;
   ctx->randa = a;
}

/* Eight-word scramble of the seeding (isaac.c:48-58), on 64-bit words. */
static void scramble(ub4 s[8]) {
   s[0] ^= s[1] << 11; s[3] += s[0]; s[1] += s[2];
   s[1] ^= s[2] >> 2;  s[4] += s[1]; s[2] += s[3];
   s[2] ^= s[3] << 8;  s[5] += s[2]; s[3] += s[4];
   s[3] ^= s[4] >> 16; s[6] += s[3]; s[4] += s[5];
   s[4] ^= s[5] << 10; s[7] += s[4]; s[5] += s[6];
   s[5] ^= s[6] >> 4;  s[0] += s[5]; s[6] += s[7];
   s[6] ^= s[7] << 8;  s[1] += s[6]; s[7] += s[0];
   s[7] ^= s[0] >> 9;  s[2] += s[7]; s[0] += s[1];
}

/* Seeding (isaac.c:61-110): with use_seed the current randrsl words are
 * the seed. */
void fl3_isaac_init(randctx *ctx, int use_seed) {
   ub4 s[8];
   int i, k;
   ctx->randa = ctx->randb = ctx->randc = 0;
   for (k = 0; k < 8; k++) s[k] = 0x9e3779b9UL;
   for (i = 0; i < 4; i++) scramble(s);
   if (use_seed) {
      int pass;
      for (pass = 0; pass < 2; pass++) {
         const ub4 *src = pass ? ctx->randmem : ctx->randrsl;
         for (i = 0; i < FLAM3_RANDSIZ; i += 8) {
            for (k = 0; k < 8; k++) s[k] += src[i + k];
            scramble(s);
            for (k = 0; k < 8; k++) ctx->randmem[i + k] = s[k];
         }
      }
   } else {
      for (i = 0; i < FLAM3_RANDSIZ; i += 8) {
         scramble(s);
         for (k = 0; k < 8; k++) ctx->randmem[i + k] = s[k];
      }
   }
   fl3_isaac_refill(ctx);
   ctx->randcnt = FLAM3_RANDSIZ;
}

/* 28-bit uniforms (flam3.c:2519-2525). */
double fl3_isaac_01(randctx *rc) {
   return (double)(fl3_irand(rc) & 0xfffffffUL) / (double)0xfffffff;
}

double fl3_isaac_11(randctx *rc) {
   long v = (long)(fl3_irand(rc) & 0xfffffffUL) - 0x7ffffffL;
   return (double)v / (double)0x7ffffff;
}
