/* LMFX fixture reader for C (the format is described in lmfx.py beside
 * this file).  Header-only: include it in one translation unit.
 *
 *   lmfx_file f;
 *   if (lmfx_open(&f, "golden_f32.lmfx") != 0) ...;
 *   lmfx_entry e;
 *   if (lmfx_find(&f, "plu_24_A", &e) == 0) ...;   e.data -> column-major payload
 *   lmfx_close(&f);
 *
 * Every entry's payload is 8-byte aligned in the loaded buffer, so a Mat's
 * data pointer (mat.h:364-368) or a Perm's (mat.h:382-385) may point at it. */
#ifndef LMFX_H_
#define LMFX_H_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { LMFX_MAT_F32 = 1, LMFX_MAT_F64 = 2, LMFX_PERM = 3, LMFX_I64 = 4, LMFX_F64 = 5 };

typedef struct {
  unsigned char *buf;
  size_t size;
  uint32_t count;
} lmfx_file;

typedef struct {
  const char *name; /* not NUL-terminated: name_len bytes */
  size_t name_len;
  int kind, rank;
  uint64_t rows, cols;
  const void *data;
} lmfx_entry;

static size_t lmfx_elem_size(int kind) { return kind == LMFX_MAT_F32 ? 4 : 8; }

static int lmfx_open(lmfx_file *f, const char *path) {
  memset(f, 0, sizeof(*f));
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  if (fseek(fp, 0, SEEK_END) != 0) return fclose(fp), -1;
  const long sz = ftell(fp);
  if (sz < 16 || fseek(fp, 0, SEEK_SET) != 0) return fclose(fp), -1;
  /* 8-byte aligned buffer: the payloads are 8-byte aligned within the file */
  f->buf = (unsigned char *)malloc(((size_t)sz + 7) / 8 * 8);
  if (!f->buf || fread(f->buf, 1, (size_t)sz, fp) != (size_t)sz) {
    fclose(fp);
    free(f->buf);
    f->buf = NULL;
    return -1;
  }
  fclose(fp);
  f->size = (size_t)sz;
  if (memcmp(f->buf, "LMATFX01", 8) != 0) {
    free(f->buf);
    f->buf = NULL;
    return -2;
  }
  memcpy(&f->count, f->buf + 8, 4);
  return 0;
}

static void lmfx_close(lmfx_file *f) {
  free(f->buf);
  f->buf = NULL;
}

/* The i-th entry (file order), or -1 past the end or on a malformed file. */
static int lmfx_entry_at(const lmfx_file *f, uint32_t i, lmfx_entry *e) {
  size_t off = 16;
  for (uint32_t k = 0; k <= i; ++k) {
    if (k >= f->count || off + 2 > f->size) return -1;
    uint16_t nl;
    memcpy(&nl, f->buf + off, 2);
    e->name = (const char *)f->buf + off + 2;
    e->name_len = nl;
    off += 2 + nl;
    off = (off + 7) / 8 * 8;
    if (off + 24 > f->size) return -1;
    e->kind = f->buf[off];
    e->rank = f->buf[off + 1];
    memcpy(&e->rows, f->buf + off + 8, 8);
    memcpy(&e->cols, f->buf + off + 16, 8);
    off += 24;
    if (e->kind < LMFX_MAT_F32 || e->kind > LMFX_F64) return -1;
    const size_t bytes = (size_t)(e->rows * e->cols) * lmfx_elem_size(e->kind);
    if (off + bytes > f->size) return -1;
    e->data = f->buf + off;
    off += (bytes + 7) / 8 * 8;
  }
  return 0;
}

static int lmfx_find(const lmfx_file *f, const char *name, lmfx_entry *e) {
  const size_t len = strlen(name);
  for (uint32_t i = 0; i < f->count; ++i) {
    if (lmfx_entry_at(f, i, e) != 0) return -1;
    if (e->name_len == len && memcmp(e->name, name, len) == 0) return 0;
  }
  return -1;
}

#endif /* LMFX_H_ */
