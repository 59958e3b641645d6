/*
 * gp_h5.c -- the MPI driver's HDF5 snapshots (DATA/<out>_<t>.h5) and their XDMF descriptions, without the HDF5
 * library.
 *
 * Layout of the reference's file (GPm/solverloop/file_writer.h:534-633, names from :895-945): one 2-D dataset of
 * native doubles per field, shape [MESH_X + 6][MESH_Y + 6] (the halo-inclusive grid, y fastest), named
 * /<phase>, /Mu_<component>, /Composition_<component> and, for non-isothermal runs, /T, in the root group.
 *
 * This file writes that content as an HDF5 file in the format's original encoding (HDF5 File Format
 * Specification, version 0 superblock): a symbol-table root group (a version 1 B-tree with one symbol-table node and
 * a local heap for the link names) and one version 1 object header per dataset with dataspace, datatype (IEEE
 * binary64 little-endian), fill-value and contiguous layout messages; the raw data follows the metadata. The
 * reader follows the same structures: superblock version 0, symbol-table groups (any B-tree depth), version 1
 * object headers with continuation messages, contiguous or compact layout, binary64 of either byte order.
 */
#include "gp_host.h"
#include "gp_snapshot.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define H5_UNDEF 0xffffffffffffffffULL
/* group B-tree widths written to the superblock: a symbol-table node holds up to 2 LEAF_K links */
#define H5_LEAF_K 32
#define H5_NODE_K 16
/* a local heap with no free block holds this marker as its free-list head (the HDF5 library's own value) */
#define H5_HEAP_FREE_NULL 1ULL

typedef struct {
  unsigned char *b;
  size_t n;
} wbuf;

static void put(wbuf *w, size_t at, const void *p, size_t n) { memcpy(w->b + at, p, n); }
static void put8(wbuf *w, size_t at, unsigned v) { w->b[at] = (unsigned char)v; }
static void put16(wbuf *w, size_t at, unsigned v) {
  for (int i = 0; i < 2; i++) w->b[at + i] = (unsigned char)(v >> (8 * i));
}
static void put32(wbuf *w, size_t at, uint32_t v) {
  for (int i = 0; i < 4; i++) w->b[at + i] = (unsigned char)(v >> (8 * i));
}
static void put64(wbuf *w, size_t at, uint64_t v) {
  for (int i = 0; i < 8; i++) w->b[at + i] = (unsigned char)(v >> (8 * i));
}
static size_t pad8(size_t n) { return (n + 7) & ~(size_t)7; }

typedef gp_field h5field;

void gp_snapshot_free(gp_field *f, int nf) {
  for (int i = 0; i < nf; i++) free(f[i].tmp);
}

int gp_snapshot_fields(const gp_input *in, const double *phi, const double *mu, const double *T_rows, gp_field *f,
                       int *nf) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int64_t rx = gp_rows_x(p), ry = gp_rows_y(p), n = rx * ry;
  int k = 0;
  for (int a = 0; a < N; a++, k++) {
    snprintf(f[k].name, sizeof f[k].name, "%s", in->phases[a]);
    f[k].data = phi + (int64_t)a * n;
    f[k].tmp = NULL;
  }
  for (int c = 0; c < NC; c++, k++) {
    snprintf(f[k].name, sizeof f[k].name, "Mu_%s", in->components[c]);
    f[k].data = mu + (int64_t)c * n;
    f[k].tmp = NULL;
  }
  double *comp = (double *)malloc(sizeof(double) * (size_t)(NC > 0 ? NC : 1) * (size_t)n);
  if (!comp) return GP_E_OOM;
  if (p->tempgrady) {
    double pc[GP_MAXP], mc[GP_MAXNC];
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) {
        const int64_t i = x * ry + y;
        for (int a = 0; a < N; a++) pc[a] = phi[(int64_t)a * n + i];
        for (int c = 0; c < NC; c++) mc[c] = mu[(int64_t)c * n + i];
        for (int c = 0; c < NC; c++) comp[(int64_t)c * n + i] = gp_composition_at(p, pc, mc, c, T_rows[y]);
      }
  } else {
    gp_composition_grid(p, phi, mu, comp, n);
  }
  for (int c = 0; c < NC; c++, k++) {
    snprintf(f[k].name, sizeof f[k].name, "Composition_%s", in->components[c]);
    f[k].data = comp + (int64_t)c * n;
    f[k].tmp = c == 0 ? comp : NULL;
  }
  if (p->tempgrady) {
    double *T = (double *)malloc(sizeof(double) * (size_t)n);
    if (!T) {
      free(comp);
      return GP_E_OOM;
    }
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) T[x * ry + y] = T_rows[y];
    snprintf(f[k].name, sizeof f[k].name, "T");
    f[k].data = T;
    f[k].tmp = T;
    k++;
  }
  if (NC == 0) free(comp);
  *nf = k;
  return GP_OK;
}

static int cmp_names(const void *a, const void *b) {
  return strcmp(((const h5field *)a)->name, ((const h5field *)b)->name);
}

int gp_write_h5(const gp_input *in, const char *path, const double *phi, const double *mu, const double *T_rows) {
  const gp_params *p = &in->p;
  if (p->tempgrady && !T_rows) return GP_E_ARG;
  const uint64_t rx = (uint64_t)gp_rows_x(p), ry = (uint64_t)gp_rows_y(p), dbytes = rx * ry * 8;
  h5field f[GP_SNAPSHOT_MAXF];
  int nf = 0, rc = gp_snapshot_fields(in, phi, mu, T_rows, f, &nf);
  if (rc) return rc;
  if (nf > 2 * H5_LEAF_K) rc = GP_E_ARG;
  /* links sorted by name, as a symbol-table node keeps them */
  h5field s[GP_SNAPSHOT_MAXF];
  memcpy(s, f, sizeof(h5field) * (size_t)nf);
  qsort(s, (size_t)nf, sizeof(h5field), cmp_names);

  /* metadata addresses */
  const size_t sb = 0, sb_size = 96;
  const size_t root_oh = sb + sb_size, root_oh_size = 16 + 8 + 16;
  const size_t btree = root_oh + root_oh_size, btree_size = 24 + (2 * H5_NODE_K + 1) * 8 + 2 * H5_NODE_K * 8;
  const size_t heap = btree + btree_size, heap_size = 32;
  size_t heap_data_size = 8, name_off[GP_SNAPSHOT_MAXF];
  for (int i = 0; i < nf; i++) {
    name_off[i] = heap_data_size;
    heap_data_size += pad8(strlen(s[i].name) + 1);
  }
  const size_t heap_data = heap + heap_size;
  const size_t snod = heap_data + heap_data_size, snod_size = 8 + 2 * H5_LEAF_K * 40;
  const size_t oh0 = snod + snod_size, oh_size = 16 + 4 * 8 + 24 + 24 + 8 + 24;
  const size_t raw0 = oh0 + (size_t)nf * oh_size;
  const uint64_t eof = raw0 + (uint64_t)nf * dbytes;
  wbuf w = {(unsigned char *)calloc(raw0, 1), raw0};
  if (!w.b || rc) {
    free(w.b);
    gp_snapshot_free(f, nf);
    return rc ? rc : GP_E_OOM;
  }

  /* superblock, version 0 */
  put(&w, sb, "\x89HDF\r\n\x1a\n", 8);
  put8(&w, sb + 13, 8); /* size of offsets */
  put8(&w, sb + 14, 8); /* size of lengths */
  put16(&w, sb + 16, H5_LEAF_K);
  put16(&w, sb + 18, H5_NODE_K);
  put64(&w, sb + 24, 0);        /* base address */
  put64(&w, sb + 32, H5_UNDEF); /* free-space info */
  put64(&w, sb + 40, eof);
  put64(&w, sb + 48, H5_UNDEF); /* driver information block */
  /* root symbol-table entry: name offset 0, its object header, cache type 1 with the B-tree and heap */
  put64(&w, sb + 56, 0);
  put64(&w, sb + 64, root_oh);
  put32(&w, sb + 72, 1);
  put64(&w, sb + 80, btree);
  put64(&w, sb + 88, heap);

  /* root object header: one symbol-table message */
  put8(&w, root_oh, 1);
  put16(&w, root_oh + 2, 1);
  put32(&w, root_oh + 4, 1);
  put32(&w, root_oh + 8, (uint32_t)(root_oh_size - 16));
  put16(&w, root_oh + 16, 0x0011);
  put16(&w, root_oh + 18, 16);
  put64(&w, root_oh + 24, btree);
  put64(&w, root_oh + 32, heap);

  /* group B-tree, one leaf level pointing at the symbol-table node; keys are heap offsets of the names bounding it */
  put(&w, btree, "TREE", 4);
  put8(&w, btree + 4, 0);
  put8(&w, btree + 5, 0);
  put16(&w, btree + 6, 1);
  put64(&w, btree + 8, H5_UNDEF);
  put64(&w, btree + 16, H5_UNDEF);
  put64(&w, btree + 24, 0);
  put64(&w, btree + 32, snod);
  put64(&w, btree + 40, nf ? name_off[nf - 1] : 0);

  /* local heap of the link names; offset 0 holds the empty name */
  put(&w, heap, "HEAP", 4);
  put64(&w, heap + 8, heap_data_size);
  put64(&w, heap + 16, H5_HEAP_FREE_NULL);
  put64(&w, heap + 24, heap_data);
  for (int i = 0; i < nf; i++) put(&w, heap_data + name_off[i], s[i].name, strlen(s[i].name));

  /* symbol-table node */
  put(&w, snod, "SNOD", 4);
  put8(&w, snod + 4, 1);
  put16(&w, snod + 6, (unsigned)nf);
  for (int i = 0; i < nf; i++) {
    const size_t e = snod + 8 + (size_t)i * 40;
    put64(&w, e, name_off[i]);
    put64(&w, e + 8, oh0 + (size_t)i * oh_size);
  }

  /* dataset object headers */
  for (int i = 0; i < nf; i++) {
    const size_t o = oh0 + (size_t)i * oh_size;
    put8(&w, o, 1);
    put16(&w, o + 2, 4);
    put32(&w, o + 4, 1);
    put32(&w, o + 8, (uint32_t)(oh_size - 16));
    size_t m = o + 16;
    /* dataspace, version 1: rank 2, current dimensions only */
    put16(&w, m, 0x0001);
    put16(&w, m + 2, 24);
    put8(&w, m + 8, 1);
    put8(&w, m + 9, 2);
    put64(&w, m + 16, rx);
    put64(&w, m + 24, ry);
    m += 8 + 24;
    /* datatype, version 1, floating point: little-endian, implied leading mantissa bit, sign bit 63; offset 0,
     * precision 64, exponent at bit 52 of 11 bits, mantissa at bit 0 of 52 bits, bias 1023 */
    put16(&w, m, 0x0003);
    put16(&w, m + 2, 24);
    put8(&w, m + 8, 0x11);
    put8(&w, m + 9, 0x20);
    put8(&w, m + 10, 63);
    put32(&w, m + 12, 8);
    put16(&w, m + 16, 0);
    put16(&w, m + 18, 64);
    put8(&w, m + 20, 52);
    put8(&w, m + 21, 11);
    put8(&w, m + 22, 0);
    put8(&w, m + 23, 52);
    put32(&w, m + 24, 1023);
    m += 8 + 24;
    /* fill value, version 2: space allocated early, fill written if set, no fill value defined */
    put16(&w, m, 0x0005);
    put16(&w, m + 2, 8);
    put8(&w, m + 8, 2);
    put8(&w, m + 9, 1);
    put8(&w, m + 10, 2);
    put8(&w, m + 11, 0);
    m += 8 + 8;
    /* layout, version 3, contiguous: address and size of the raw data */
    put16(&w, m, 0x0008);
    put16(&w, m + 2, 24);
    put8(&w, m + 8, 3);
    put8(&w, m + 9, 1);
    put64(&w, m + 10, raw0 + (uint64_t)i * dbytes);
    put64(&w, m + 18, dbytes);
  }

  FILE *fp = fopen(path, "wb");
  rc = fp ? GP_OK : GP_E_IO;
  if (fp && fwrite(w.b, 1, w.n, fp) != w.n) rc = GP_E_IO;
  /* raw data: little-endian doubles, in the order of the sorted links */
  unsigned char *row = (unsigned char *)malloc((size_t)ry * 8);
  for (int i = 0; fp && !rc && i < nf; i++)
    for (uint64_t x = 0; x < rx && !rc; x++) {
      for (uint64_t y = 0; y < ry; y++) {
        uint64_t bits;
        memcpy(&bits, &s[i].data[x * ry + y], 8);
        for (int b = 0; b < 8; b++) row[y * 8 + b] = (unsigned char)(bits >> (8 * b));
      }
      if (fwrite(row, 1, (size_t)ry * 8, fp) != (size_t)ry * 8) rc = GP_E_IO;
    }
  free(row);
  if (fp && fclose(fp) != 0) rc = GP_E_IO;
  free(w.b);
  gp_snapshot_free(f, nf);
  return rc;
}

/* ------------------------------------------------------------------------------------------------ reader */
typedef struct {
  FILE *fp;
  uint64_t base;
  int so, sl; /* sizes of offsets and lengths */
} h5r;

static int rd(h5r *r, uint64_t at, void *dst, size_t n) {
  if (fseeko(r->fp, (off_t)(r->base + at), SEEK_SET) != 0) return GP_E_IO;
  return fread(dst, 1, n, r->fp) == n ? GP_OK : GP_E_IO;
}
static uint64_t le(const unsigned char *b, int n) {
  uint64_t v = 0;
  for (int i = n - 1; i >= 0; i--) v = (v << 8) | b[i];
  return v;
}

typedef struct {
  uint64_t dims[2];
  int rank, be, ok_type;
  uint64_t addr, size;
  int compact;
  unsigned char *compact_data;
} h5dset;

/* the messages of a version 1 object header, following continuation messages */
static int read_object(h5r *r, uint64_t oh, h5dset *d) {
  unsigned char pre[16];
  if (rd(r, oh, pre, 16)) return GP_E_IO;
  if (pre[0] != 1) return GP_E_ARG;
  uint64_t block = oh + 16, len = le(pre + 8, 4);
  int nleft = (int)le(pre + 2, 2);
  uint64_t conts[64][2];
  int ncont = 0, ci = 0;
  while (nleft > 0) {
    unsigned char *buf = (unsigned char *)malloc((size_t)len);
    if (!buf || rd(r, block, buf, (size_t)len)) {
      free(buf);
      return GP_E_IO;
    }
    for (uint64_t m = 0; m + 8 <= len && nleft > 0; nleft--) {
      const unsigned type = (unsigned)le(buf + m, 2), sz = (unsigned)le(buf + m + 2, 2);
      const unsigned char *b = buf + m + 8;
      if (m + 8 + sz > len) break;
      if (type == 0x0001) {
        d->rank = b[1];
        const int off = b[0] == 1 ? 8 : 4;
        for (int k = 0; k < d->rank && k < 2; k++) d->dims[k] = le(b + off + k * r->sl, r->sl);
      } else if (type == 0x0003) {
        d->ok_type = (b[0] & 0x0f) == 1 && le(b + 4, 4) == 8;
        d->be = b[1] & 1;
      } else if (type == 0x0008) {
        if (b[0] == 3 && b[1] == 1) {
          d->addr = le(b + 2, r->so);
          d->size = le(b + 2 + r->so, r->sl);
        } else if (b[0] == 3 && b[1] == 0) {
          d->size = le(b + 2, 2);
          d->compact = 1;
          free(d->compact_data);
          d->compact_data = (unsigned char *)malloc((size_t)d->size);
          if (d->compact_data) memcpy(d->compact_data, b + 4, (size_t)d->size);
        }
      } else if (type == 0x0010 && ncont < 64) {
        conts[ncont][0] = le(b, r->so);
        conts[ncont][1] = le(b + r->so, r->sl);
        ncont++;
      }
      m += 8 + sz;
    }
    free(buf);
    if (nleft <= 0 || ci >= ncont) break;
    block = conts[ci][0];
    len = conts[ci][1];
    ci++;
  }
  return GP_OK;
}

typedef struct {
  const char *want;
  uint64_t found;
} h5look;

/* the object header address of a link of the root group: a walk of its B-tree (any depth) */
static int find_link(h5r *r, uint64_t node, uint64_t heap_data, int leafk, h5look *lk) {
  unsigned char hdr[24];
  if (rd(r, node, hdr, 8)) return GP_E_IO;
  if (!memcmp(hdr, "SNOD", 4)) {
    const int n = (int)le(hdr + 6, 2);
    const size_t esz = (size_t)(2 * r->so + 24);
    unsigned char *e = (unsigned char *)malloc(esz * (size_t)(n > 0 ? n : 1));
    if (!e || rd(r, node + 8, e, esz * (size_t)n)) {
      free(e);
      return GP_E_IO;
    }
    for (int i = 0; i < n && !lk->found; i++) {
      char name[256] = {0};
      if (rd(r, heap_data + le(e + esz * i, r->so), name, sizeof name - 1) == GP_OK && !strcmp(name, lk->want))
        lk->found = le(e + esz * i + r->so, r->so);
    }
    free(e);
    return GP_OK;
  }
  if (memcmp(hdr, "TREE", 4) || hdr[4] != 0) return GP_E_ARG;
  if (rd(r, node, hdr, 8)) return GP_E_IO;
  const int used = (int)le(hdr + 6, 2);
  const uint64_t body = node + 8 + 2 * (uint64_t)r->so;
  for (int i = 0; i < used && !lk->found; i++) {
    unsigned char c[8];
    if (rd(r, body + (uint64_t)r->sl + (uint64_t)i * (r->sl + r->so), c, (size_t)r->so)) return GP_E_IO;
    const int rc = find_link(r, le(c, r->so), heap_data, leafk, lk);
    if (rc) return rc;
  }
  return GP_OK;
}

static int read_dataset(h5r *r, uint64_t btree, uint64_t heap_data, int leafk, const char *name, uint64_t rx,
                        uint64_t ry, double *out) {
  h5look lk = {name, 0};
  int rc = find_link(r, btree, heap_data, leafk, &lk);
  if (rc) return rc;
  if (!lk.found) return GP_E_ARG;
  h5dset d;
  memset(&d, 0, sizeof d);
  if ((rc = read_object(r, lk.found, &d))) return rc;
  if (d.rank != 2 || d.dims[0] != rx || d.dims[1] != ry || !d.ok_type || d.size != rx * ry * 8) rc = GP_E_ARG;
  unsigned char *raw = (unsigned char *)malloc((size_t)(rx * ry * 8));
  if (!rc && !raw) rc = GP_E_OOM;
  if (!rc) {
    if (d.compact) memcpy(raw, d.compact_data, (size_t)d.size);
    else rc = rd(r, d.addr, raw, (size_t)d.size);
  }
  for (uint64_t i = 0; !rc && i < rx * ry; i++) {
    uint64_t bits = 0;
    for (int b = 0; b < 8; b++) bits |= (uint64_t)raw[i * 8 + (d.be ? 7 - b : b)] << (8 * b);
    memcpy(&out[i], &bits, 8);
  }
  free(raw);
  free(d.compact_data);
  return rc;
}

int gp_read_h5(const gp_input *in, const char *path, double *phi, double *mu) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const uint64_t rx = (uint64_t)gp_rows_x(p), ry = (uint64_t)gp_rows_y(p);
  h5r r = {fopen(path, "rb"), 0, 8, 8};
  if (!r.fp) return GP_E_IO;
  unsigned char sb[96];
  int rc = rd(&r, 0, sb, 96);
  if (!rc && (memcmp(sb, "\x89HDF\r\n\x1a\n", 8) || sb[8] != 0)) rc = GP_E_ARG;
  uint64_t btree = 0, heap_data = 0;
  int leafk = 0;
  if (!rc) {
    r.so = sb[13];
    r.sl = sb[14];
    if (r.so != 8 || r.sl != 8) rc = GP_E_ARG;
    leafk = (int)le(sb + 16, 2);
    r.base = le(sb + 24, 8);
  }
  if (!rc) {
    /* the root entry's scratch pad holds the B-tree and heap when its cache type is 1; else read its header */
    const unsigned char *e = sb + 56;
    uint64_t heap = 0;
    if (le(e + 16, 4) == 1) {
      btree = le(e + 24, 8);
      heap = le(e + 32, 8);
    } else {
      unsigned char pre[16], msg[40];
      const uint64_t oh = le(e + 8, 8);
      rc = rd(&r, oh, pre, 16);
      if (!rc) rc = rd(&r, oh + 16, msg, 40);
      if (!rc && le(msg, 2) == 0x0011) {
        btree = le(msg + 8, 8);
        heap = le(msg + 16, 8);
      } else if (!rc) {
        rc = GP_E_ARG;
      }
    }
    unsigned char hh[32];
    if (!rc) rc = rd(&r, heap, hh, 32);
    if (!rc && memcmp(hh, "HEAP", 4)) rc = GP_E_ARG;
    if (!rc) heap_data = le(hh + 24, 8);
  }
  char name[GP_NAMELEN + 16];
  for (int a = 0; !rc && a < N; a++) {
    snprintf(name, sizeof name, "%s", in->phases[a]);
    rc = read_dataset(&r, btree, heap_data, leafk, name, rx, ry, phi + (int64_t)a * (int64_t)(rx * ry));
  }
  for (int c = 0; !rc && c < NC; c++) {
    snprintf(name, sizeof name, "Mu_%s", in->components[c]);
    rc = read_dataset(&r, btree, heap_data, leafk, name, rx, ry, mu + (int64_t)c * (int64_t)(rx * ry));
  }
  fclose(r.fp);
  return rc;
}

/* write_xdmf_xml (GPm/write_xdmf.c:66-146), 2-D: one Attribute per dataset of <h5_name> */
int gp_write_xdmf(const gp_input *in, const char *path, const char *h5_name) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const long MX = (long)p->mesh_x + 6, MY = (long)p->mesh_y + 6;
  FILE *x = fopen(path, "w");
  if (!x) return GP_E_IO;
  fprintf(x, "<?xml version=\"1.0\" ?>\n");
  fprintf(x, "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
  fprintf(x, "<Xdmf Version=\"2.0\">\n");
  fprintf(x, " <Domain>\n");
  fprintf(x, "   <Grid Name=\"mesh1\" GridType=\"Uniform\">\n");
  fprintf(x, "     <Topology TopologyType=\"2DCoRectMesh\" NumberOfElements=\"%ld %ld\"/>\n", MX, MY);
  fprintf(x, "     <Geometry GeometryType=\"ORIGIN_DXDY\">\n");
  fprintf(x, "       <DataItem Dimensions=\"%d\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n", 2);
  fprintf(x, "       \t0 0\n");
  fprintf(x, "       </DataItem>\n");
  fprintf(x, "       <DataItem Dimensions=\"%d\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n", 2);
  fprintf(x, "       \t%lf %lf\n", 1.0, 1.0);
  fprintf(x, "       </DataItem>\n");
  fprintf(x, "     </Geometry>\n");
  char attr[GP_NAMELEN + 16];
  const int nattr = N + NC + (in->writecomposition ? NC : 0) + (p->tempgrady ? 1 : 0);
  for (int i = 0; i < nattr; i++) {
    if (i < N) snprintf(attr, sizeof attr, "%s", in->phases[i]);
    else if (i < N + NC) snprintf(attr, sizeof attr, "Mu_%s", in->components[i - N]);
    else if (in->writecomposition && i < N + 2 * NC) snprintf(attr, sizeof attr, "Composition_%s", in->components[i - N - NC]);
    else snprintf(attr, sizeof attr, "T");
    fprintf(x, "     <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Node\">\n", attr);
    fprintf(x, "       <DataItem Dimensions=\"%ld %ld\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n", MX, MY);
    fprintf(x, "        %s:/%s\n", h5_name, attr);
    fprintf(x, "       </DataItem>\n");
    fprintf(x, "     </Attribute>\n");
  }
  fprintf(x, "   </Grid>\n");
  fprintf(x, " </Domain>\n");
  fprintf(x, "</Xdmf>\n");
  return fclose(x) == 0 ? GP_OK : GP_E_IO;
}
