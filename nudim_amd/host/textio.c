/* textio.c -- threaded text I/O for the command-line tools (see textio.h). */
#include "textio.h"

#include <ctype.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int textio_threads(void) {
  const char *e = getenv("NUDIM_IO_THREADS");
  long n = e ? strtol(e, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
  if (n < 1) n = 1;
  if (n > 16) n = 16;
  return (int)n;
}

static char *slurp(const char *path, size_t *len) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return NULL;
  size_t cap = 1 << 20, n = 0;
  char *buf = (char *)malloc(cap + 1);
  if (!buf) {
    fclose(fp);
    return NULL;
  }
  for (;;) {
    size_t got = fread(buf + n, 1, cap - n, fp);
    n += got;
    if (n < cap) break;
    cap *= 2;
    char *nb = (char *)realloc(buf, cap + 1);
    if (!nb) {
      free(buf);
      fclose(fp);
      return NULL;
    }
    buf = nb;
  }
  fclose(fp);
  buf[n] = '\0';
  *len = n;
  return buf;
}

typedef struct {
  const char *b, *e;  /* segment [b, e), starts and ends at a token boundary */
  double *out;        /* NULL: count only */
  long limit, n;
  int bad;            /* a token strtod rejects */
} seg_t;

static void *parse_seg(void *arg) {
  seg_t *s = (seg_t *)arg;
  const char *p = s->b;
  s->n = 0;
  s->bad = 0;
  while (p < s->e && s->n < s->limit) {
    while (p < s->e && isspace((unsigned char)*p)) p++;
    if (p >= s->e) break;
    if (s->out) {
      char *q;
      double v = strtod(p, &q);
      if (q == p) {
        s->bad = 1;
        break;
      }
      s->out[s->n] = v;
      p = q;
    } else {
      while (p < s->e && !isspace((unsigned char)*p)) p++;
    }
    s->n++;
  }
  return NULL;
}

long textio_read_doubles(const char *path, double *out, long count, int nthreads) {
  size_t len = 0;
  char *buf = slurp(path, &len);
  if (!buf) return -1;
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > len / 4096 + 1) nthreads = (int)(len / 4096 + 1);
  seg_t *sg = (seg_t *)calloc(nthreads, sizeof(seg_t));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  /* cut at whitespace so that no token straddles two segments */
  const char *start = buf, *end = buf + len;
  for (int t = 0; t < nthreads; t++) {
    const char *e = (t == nthreads - 1) ? end : buf + (len * (t + 1)) / nthreads;
    if (e < start) e = start;
    while (e < end && !isspace((unsigned char)*e)) e++;
    sg[t].b = start;
    sg[t].e = e;
    start = e;
  }
  /* pass 1: token counts; pass 2: parse into place */
  for (int t = 0; t < nthreads; t++) {
    sg[t].out = NULL;
    sg[t].limit = count;
    pthread_create(&th[t], NULL, parse_seg, &sg[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  long off = 0;
  for (int t = 0; t < nthreads; t++) {
    long take = sg[t].n;
    if (off + take > count) take = count - off;
    sg[t].out = out + off;
    sg[t].limit = take;
    off += take;
  }
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, parse_seg, &sg[t]);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  /* values up to the first rejected token, as a sequential fscanf would */
  long total = 0;
  for (int t = 0; t < nthreads; t++) {
    total += sg[t].n;
    if (sg[t].bad || sg[t].n < sg[t].limit) break;
  }
  free(sg);
  free(th);
  free(buf);
  return total;
}

typedef struct {
  const double *v;
  long n;
  const char *fmt;
  char *buf;
  size_t len;
  int err;
} fmt_t;

static void *format_seg(void *arg) {
  fmt_t *f = (fmt_t *)arg;
  size_t cap = (size_t)f->n * 16 + 64, len = 0;
  char *buf = (char *)malloc(cap);
  f->err = buf == NULL;
  for (long i = 0; i < f->n && !f->err; i++) {
    for (;;) {
      int w = snprintf(buf + len, cap - len, f->fmt, f->v[i]);
      if (w < 0) {
        f->err = 1;
        break;
      }
      if ((size_t)w < cap - len) {
        len += (size_t)w;
        break;
      }
      cap *= 2;
      char *nb = (char *)realloc(buf, cap);
      if (!nb) {
        f->err = 1;
        break;
      }
      buf = nb;
    }
  }
  f->buf = buf;
  f->len = len;
  return NULL;
}

int textio_write_doubles(const char *path, const double *v, long count, const char *fmt, int nthreads) {
  FILE *fp = fopen(path, "w");
  if (!fp) return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > count / 1024 + 1) nthreads = (int)(count / 1024 + 1);
  fmt_t *fs = (fmt_t *)calloc(nthreads, sizeof(fmt_t));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  for (int t = 0; t < nthreads; t++) {
    long a = count * t / nthreads, b = count * (t + 1) / nthreads;
    fs[t].v = v + a;
    fs[t].n = b - a;
    fs[t].fmt = fmt;
    pthread_create(&th[t], NULL, format_seg, &fs[t]);
  }
  int rc = 0;
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads; t++) {
    if (fs[t].err || fwrite(fs[t].buf, 1, fs[t].len, fp) != fs[t].len) rc = -1;
    free(fs[t].buf);
  }
  if (fclose(fp) != 0) rc = -1;
  free(fs);
  free(th);
  return rc;
}
